/*
 * kestrel_amd.h -- C ABI of the MI355X (gfx950) pairwise Jacquard / kinship
 * maximum-likelihood fit.  Plain C types only; every call returns 0 on success
 * or a negative KESTREL_E* code, with a message in kestrel_last_error().
 *
 * Reference interfaces replaced (paths in the kestrel 0.1.2 source tree):
 *   kestrel_cohort_create     wrapper.rs:89-207 up to the pair loop: allele
 *                             frequencies by per-site SQP (allele/mod.rs:37-134,
 *                             maf default 1/(2N) wrapper.rs:101), zero-frequency
 *                             allele removal and non-segregating site drop
 *                             (wrapper.rs:237-410), lane layout
 *                             (lanearray2.rs:48-80); input = hard calls as the
 *                             one-hot GL rows of vcf.rs:338-382.
 *   kestrel_fit_pairs         coefficients.rs:40-141
 *                             calculate_relatedness_coefficients_gl: every pair
 *                             (x <= y) fitted by sqp.rs:49-93; Output records of
 *                             coefficients.rs:143-151.
 *   kestrel_evaluate          the operator boundary sqp.rs:8-14 implemented by
 *                             JacquardCoefficientData (jacquard/mod.rs:6-21):
 *                             objective + gradient + Hessian at given points.
 *   kestrel_tune_default      sqp.rs:31-46 Tuneables::new(iterations, 10)
 *                             (coefficients.rs:73).
 *
 * Ownership: the caller owns every host buffer; a cohort owns its device
 * copies.  A cohort is used from one host thread at a time.
 */
#ifndef KESTREL_AMD_H
#define KESTREL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KESTREL_OK 0
#define KESTREL_EINVAL (-1)      /* bad argument (reference: assert!/panic!) */
#define KESTREL_EDEVICE (-2)     /* HIP runtime error or no gfx950 device */
#define KESTREL_ENOMEM (-3)      /* device allocation failed */
#define KESTREL_EDATA (-4)       /* a site has no called sample (lanevector.rs:73-76) */

/* per-pair status in kestrel_output.status */
#define KESTREL_PAIR_OK 0u
#define KESTREL_PAIR_NO_SITES 1u /* no usable site: the reference panics (lanevector.rs:133-138) */

/* precision forms: KESTREL_FP64 is the reference's arithmetic (bit-exact in
 * the device's summation order).  KESTREL_MIXED keeps in FP32 the c_s table
 * (or, for likelihood rows, the c_s construction), 1/L (v_rcp_f32, about one
 * FP32 ulp), the scaled terms c/L and the Hessian sums over a segment of
 * sites; L = eps + c.D, the log and the objective are FP64; each gradient
 * term is the exact FP64 product of the FP32 c and the FP32 1/L, accumulated
 * in FP64, so the KKT residual of the convergence test carries FP32 per-site
 * error.  The QP, the line search and the step are FP64.  The condition of a
 * converged pair whose estimate from the FP32-summed Hessian reaches 100 is
 * computed again from an FP64 evaluation at its final point, so a
 * non-identifiable pair keeps its large (or infinite) condition.  Bound:
 * with u = 2^-24 and segments of n sites each Hessian term is within
 * (8 + n) u sqrt(H_ii H_jj) of its FP64 value (||dH|| <= 9 (8 + n) u ||H||,
 * 2.2e-3 at n = 4096), enough for a log10 condition below 100 to move less
 * than 0.1; the converged point is fixed by the FP64-accumulated gradient's
 * test, and measured records differ from the FP64 form's by under 1e-7 in
 * Delta on well-conditioned pairs.  2-bit cohorts take it at
 * kestrel_cohort_create, others through the option "precision". */
#define KESTREL_FP64 0
#define KESTREL_MIXED 1

/* maf argument of the cohort constructors: exactly this value selects the
 * reference's default 1/(2 n_samples) (wrapper.rs:101); any other value must
 * lie within 0 to 1 (wrapper.rs:101-106 asserts it), else KESTREL_EINVAL. */
#define KESTREL_MAF_DEFAULT (-1.0)

/* sqp.rs:16-29 Tuneables */
typedef struct {
    uint64_t sqp_max_iter;
    double sqp_conv_tol;
    uint64_t qp_max_iter;
    double qp_conv_tol;
    double qp_zero_search_tol;
    uint64_t bls_max_iter;
    double bls_sufficient_decrease;
    double bls_step_size_reduce;
    double epsilon;
} kestrel_tune;

/* coefficients.rs:143-151 Output + usable-site count n_xy + status (112 bytes) */
typedef struct {
    uint32_t x, y;
    double jacquard[9];
    double obj;
    uint64_t iters;
    double cond;
    uint32_t n_sites;
    uint32_t status;
} kestrel_output;

typedef struct kestrel_cohort kestrel_cohort;

/* Defaults of sqp.rs:31-46 with sqp_max_iter = iterations (cli.rs:33-35) and
 * qp_max_iter = 10 (coefficients.rs:73). */
int kestrel_tune_default(kestrel_tune* tune, uint64_t iterations);

/* Upload a cohort of 2-bit genotype codes and prepare it for pair fits.
 *   packed      n_samples rows of row_stride_bytes; site s of a row is in bits
 *               2*(s%4) of byte s/4; code 0 = 0/0, 1 = 0/1, 2 = 1/1 (the
 *               unphased VCF index, coefficients.rs:18-38), 3 = missing.
 *   af          NULL: estimate allele frequencies on the device as the
 *               reference does for -f GL/PL (allele/mod.rs:37-134) with
 *               af_iterations SQP iterations and minimum allele frequency maf
 *               (KESTREL_MAF_DEFAULT: 1/(2 n_samples), wrapper.rs:101).  Non-NULL: n_sites
 *               x 2 frequencies (REF, ALT) used as given.
 * Sites where either frequency is 0 are dropped (wrapper.rs:352-410).
 * Returns KESTREL_EDATA if some site has no called sample. */
int kestrel_cohort_create(kestrel_cohort** out, const uint8_t* packed, uint32_t n_samples, uint32_t n_sites,
                          size_t row_stride_bytes, const double* af, double maf, uint64_t af_iterations,
                          int precision);

/* A cohort of soft genotype likelihoods for the maximum-likelihood fit.
 * gl: n_sites x n_samples x 3 doubles, site-major as the reference's
 * GenotypeLikelihoods (vcf.rs:321-323), each triple the linear-scale
 * likelihoods of 0/0, 0/1, 1/1 (10^(gl - max gl) or 10^(-pl/10), vcf.rs:368,
 * 407); a triple with a NaN is a missing call.  Replaces wrapper.rs:89-207 up
 * to the pair loop: allele frequencies by the per-site SQP on these rows
 * (allele/mod.rs:37-134, KESTREL_MAF_DEFAULT = 1/(2N)) unless af (n_sites x 2) is given,
 * kept sites as kestrel_cohort_create.  kestrel_fit_pairs / kestrel_evaluate
 * then build c_s per pair and site as mixture_row2 (mixture.rs:96-236).
 * Biallelic sites only.  When every called row is one-hot (an exact 1.0
 * beside two zeros: hard calls fed as likelihoods) the cohort is held as its
 * 2-bit codes and fitted as a kestrel_cohort_create cohort (c_s of a one-hot
 * row is m(gx, gy) bit for bit, so the records are the same; it can then
 * serve).  kestrel_cohort_create_gl_multi does the same for its 2-allele
 * part. */
int kestrel_cohort_create_gl(kestrel_cohort** out, const double* gl, uint32_t n_samples, uint32_t n_sites,
                             const double* af, double maf, uint64_t af_iterations);

/* Soft likelihoods of sites with 2, 3 and 4 alleles, the buckets of the
 * reference's GenotypeLikelihoods {gl2, gl3, gl4} (vcf.rs:270-330):
 * glA is nA x n_samples x G doubles (G = 3, 6, 10 likelihoods in VCF genotype
 * order), site-major, a NaN marking a missing call; NULL when nA = 0.
 * Replaces wrapper.rs:89-130 + split_gl_af (wrapper.rs:237-410): allele
 * frequencies per site by the A-allele SQP (allele/mod.rs:37-134, QP cap
 * A + 1, KESTREL_MAF_DEFAULT = 1/(2N)), alleles of frequency 0 removed and every site
 * re-bucketed by its remaining allele count (sites left with one allele are
 * dropped).  Pair fits then add the 3- and 4-allele sites' c_s from the generic
 * mixture over IIS_3 / IIS_4 (mixture.rs:8-60) after the 2-allele sites. */
int kestrel_cohort_create_gl_multi(kestrel_cohort** out, uint32_t n_samples, const double* gl2, uint32_t n2,
                                   const double* gl3, uint32_t n3, const double* gl4, uint32_t n4, double maf,
                                   uint64_t af_iterations);
/* Of a kestrel_cohort_create_gl_multi cohort: part_sites = sites with 2, 3, 4
 * alleles after zero-frequency removal; afA (nA x A, may be NULL) = the
 * frequencies fitted for the input buckets.  Of a
 * kestrel_cohort_create_gt_multi cohort: part_sites = kept sites of each
 * bucket; afA = the frequencies of the input sites (NaN for a dropped one). */
int kestrel_cohort_multi_info(const kestrel_cohort* c, uint32_t part_sites[3], double* af2, double* af3,
                              double* af4);

/* Frequencies of all input sites (n_sites x 2) and the kept-site mask. */
int kestrel_cohort_allele_freqs(const kestrel_cohort* c, double* af_out, uint8_t* kept_out);
/* Number of kept (segregating) sites. */
int kestrel_cohort_n_kept(const kestrel_cohort* c, uint32_t* n_kept);

/* Fit pairs.  pairs: n_pairs (x, y) sample indices, or NULL for the upper
 * triangle including the diagonal of the first n_select samples in row-major
 * order (coefficients.rs:57,87), n_pairs = n_select (n_select + 1) / 2.
 * out has one record per pair, in pair order. Blocks until done. */
int kestrel_fit_pairs(kestrel_cohort* c, const kestrel_tune* tune, const uint32_t* pairs, size_t n_pairs,
                      uint32_t n_select, kestrel_output* out);

/* Objective, gradient and Hessian of the Jacquard likelihood at n points:
 * pairs n x 2, points n x 9; f n, g n x 9, h n x 81 (row-major), n_sites n.
 * Any point and any eps, as the reference's Objective / GradHess accept: a
 * point with finite, non-negative coefficients, eps in [1e-18, 1e14] and
 * eps + sum(D) x (largest likelihood)^2 <= 1e15 runs on the sweeps (every
 * L = eps + c.D is then a positive normal number, as are 1/L and the products
 * of 16 L the objective takes logs of); any other point is evaluated with
 * IEEE division and per-site logs wherever a 16-site group holds an L that is
 * not a positive normal number, so L = 0 gives f = +inf, a negative L gives
 * NaN and 1/L is the IEEE quotient (log/scalar.rs:49-57, grad_hess.rs:122),
 * in FP64 whatever the cohort's precision form.  kestrel_fit_pairs requires
 * eps in [1e-18, 1e14] (its points lie on the simplex). */
int kestrel_evaluate(kestrel_cohort* c, const uint32_t* pairs, const double* points, size_t n, double eps,
                     double* f, double* g, double* h, uint32_t* n_sites);

/* kestrel_evaluate for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference.  pairs (n x 2), points (n x 9), f (n),
 * g (n x 9), h (n x 81) and n_sites (n) are DEVICE pointers; results are
 * kestrel_evaluate's bit for bit.  The sweep's slot arrays, segment
 * partials and (class-sorted sweep, the default for FP64 2-bit cohorts) each
 * point's pattern lists live in a fixed 64 MiB array of device memory that
 * each GPU the library runs on holds until the process ends, so a call takes
 * at most 64 MiB / (88 + 448 x segments + lists) bytes' worth of points
 * (segments: one per 4,096 kept sites by default; lists: 8 + 18,960 bytes
 * per 8,192 kept sites with the class-sorted sweep; about 5,900 points at
 * 100,000 sites with the table sweep, 256 with the class-sorted one)
 * and returns KESTREL_EINVAL beyond that.  The call allocates nothing, runs on
 * the cohort's stream and returns when the results are written.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (from several
 * threads or cohorts) race. */
int kestrel_evaluate_noalloc(kestrel_cohort* c, const uint32_t* pairs, const double* points, size_t n, double eps,
                             double* f, double* g, double* h, uint32_t* n_sites);

/* The CLS / Brier estimator of the CLI default -f GT.
 *   kestrel_cohort_create_gt  replaces wrapper.rs:22-87 up to the pair loop:
 *                             sites with a missing call in any sample are
 *                             dropped (vcf.rs:217-221), frequencies are
 *                             counted (wrapper.rs:412-428) unless given
 *                             (n_sites x 2, -a info), every other site is kept.
 *   kestrel_fit_pairs_gt      replaces coefficients.rs:166-278
 *                             calculate_coefficients_gt: one 9x9 Gram matrix Q
 *                             (quadratic.rs, coefficients.rs:280-320), per pair
 *                             c = -(1/V) sum_s m_s[gx][gy] (cls.rs:28-49) in site
 *                             order, the QP (2 variables for a self pair), brier
 *                             in obj, QP iterations in iters (cap = iterations).
 *                             q_out (81 doubles, may be NULL) receives Q. */
int kestrel_cohort_create_gt(kestrel_cohort** out, const uint8_t* packed, uint32_t n_samples, uint32_t n_sites,
                             size_t row_stride_bytes, const double* af);
int kestrel_fit_pairs_gt(kestrel_cohort* c, uint64_t iterations, const uint32_t* pairs, size_t n_pairs,
                         uint32_t n_select, kestrel_output* out, double* q_out);

/* kestrel_cohort_create_gt with sites of 3 and 4 alleles: replaces
 * vcf.rs:95-191 parse_vcf_gt's three buckets and wrapper.rs:22-87 up to the
 * pair loop.  packed2 / n2 / row_stride_bytes: the 2-allele sites as for
 * kestrel_cohort_create_gt; gt3 (n_samples x n3) and gt4 (n_samples x n4):
 * genotype indices of the 3- and 4-allele sites, sample-major, VCF order
 * (j(j+1)/2 + i for alleles i <= j, coefficients.rs:19-38), 255 = a missing
 * call (the site is dropped, vcf.rs:217-221); any bucket may be empty.
 * af2 / af3 / af4 (n_A x A, reference allele first; -a info) are all given or
 * all NULL (counted per bucket, wrapper.rs:36-40).  kestrel_fit_pairs_gt then
 * fits the reference's calculate_coefficients_gt: Q from the m rows of the
 * 2-, 3- and 4-allele sites in that order, c = (c2 + (c3 + c4)) / -V over all
 * V kept sites.  kestrel_cohort_multi_info reports the kept sites of each
 * bucket and its frequencies (input sites, NaN for a dropped one). */
int kestrel_cohort_create_gt_multi(kestrel_cohort** out, uint32_t n_samples, const uint8_t* packed2, uint32_t n2,
                                   size_t row_stride_bytes, const uint8_t* gt3, uint32_t n3, const uint8_t* gt4,
                                   uint32_t n4, const double* af2, const double* af3, const double* af4);

/* Tuning knobs: "segment_sites" (sites per segment of a sweep, a multiple of
 * 32, default 4096; fixes the summation order, so it is a cohort property),
 * "slots" (pairs in flight per slot group, 0 = fill the device), "streams"
 * (1 or 2 slot groups on separate HIP streams, default 2: one group's fold and
 * step run beside the other's sweep; results do not change), "batch" (sweep
 * and step iterations queued per host synchronisation, 1 to 64, default 8;
 * results do not change), "graph" (1: a fit captures each slot group's batch
 * of sweeps and steps once into a HIP graph and relaunches it; timing stats
 * are then per batch; results do not change), "pair_tile" (claim order: pairs
 * are taken in tiles of this many samples so that few sample rows are in
 * flight; 0 = list order, -1 = automatic: sqrt(pairs in flight), within 32
 * to 256, for likelihood-row cohorts, list order for 2-bit ones; records stay
 * in list order and do not change), "skip_masked" (2-bit cohorts: the
 * 256-slot workgroups of a fit claim pairs in runs of consecutive claims that
 * share their first sample, and a sweep passes over the sites where a
 * wavefront's shared sample is missing, which add nothing; 1 = on, 0 = off, -1 = on when
 * at least 2 % of the kept calls are missing and the form is FP64 (default;
 * the table sweep only: the class-sorted sweep's lists hold no missing
 * site); results do not change), "pair_lanes" (FP64 fits of 2-bit and biallelic
 * likelihood-row cohorts: 1 = one
 * pair per lane (default), 64 = one pair per wavefront with its sites across
 * the 64 lanes and the lanes' sums added in a butterfly at each segment's
 * end, for fits of fewer pairs than the device has lanes; -1 = 64 for fits
 * of at most 4,096 pairs, 1 otherwise; records are those of another
 * summation order, within the reference tolerances of the default form's),
 * "class_sweep" (FP64 fits, evaluations and serving of 2-bit cohorts: 1 =
 * the class-sorted sweep (default): one pair per wavefront, its usable sites
 * sorted once per pair into lists by genotype-pair pattern and summed in each
 * pattern's basis of at most four values (mixture.rs:137-227), a third of
 * the table sweep's per-site arithmetic; 0 = the table sweep of one pair per
 * lane; the two give the reference's sums in different summation orders, so
 * records differ in the last bits), "class_waves" (pairs per workgroup of
 * the class-sorted sweep: 8, 12, 14 or 16 (default); results do not change),
 * "class_list_mib" (device memory the class-sorted sweep's pattern lists,
 * about 2 bytes per kept site and slot, may take in one fit, evaluation or
 * serving engine, in MiB; 0 = a quarter of the free memory, at most 16 GiB
 * (default); fewer slots fit, evaluations go in chunks of the slots that
 * fit, results do not change; a call returns KESTREL_EDEVICE when 256 slots
 * a group do not fit),
 * "precision" (KESTREL_FP64 or KESTREL_MIXED, for 2-bit and likelihood-row
 * cohorts, tri/tetra-allelic parts included: the form of the next fits; this
 * one changes records),
 * "serve_policy" (order in
 * which a serving engine started afterwards feeds admitted requests to the
 * device: 1 = shortest remaining first, aged by the device's pair rate times
 * the time since submission so that large requests are not starved (default),
 * 0 = round robin). */
int kestrel_set_option(kestrel_cohort* c, const char* key, int64_t value);

/* Launch-shape autotuner (no counterpart in the reference; the reference's
 * thread pool has no shape to tune).  Fits the sample `pairs` (n_pairs x 2, a
 * random sample of the job, at least a few times the device's pairs in
 * flight) twice with each candidate shape -- "slots" at 1/8 to 3 device
 * fills, "streams" 1 and 2, "batch" 4, 8, 16; a shape only when the sample
 * holds 4 times its pairs in flight -- and predicts the time of a job of
 * job_pairs pairs (0 = n_pairs) from the sample's steady rate (pairs done
 * while every slot was busy) plus its drain and host time.  Keeps the
 * fastest prediction and sets those options on the cohort.  None of them
 * changes a fit's results. */
typedef struct {
    int64_t slots, streams, batch;  /* the shape chosen (slots 0 = fill the device) */
    double pairs_per_s;             /* its predicted job throughput */
    double default_pairs_per_s;     /* that of the shape in force before */
    uint32_t trials;                /* sample fits run */
} kestrel_autotune_result;
int kestrel_autotune(kestrel_cohort* c, const kestrel_tune* tune, const uint32_t* pairs, size_t n_pairs,
                     uint64_t job_pairs, kestrel_autotune_result* res);

/* Timing of the last kestrel_fit_pairs: sweep-kernel milliseconds (sum of
 * launch durations by HIP events), number of sweeps, wall milliseconds.  After
 * kestrel_fit_pairs_gt: the milliseconds of its per-pair kernels, 1, wall. */
int kestrel_last_fit_stats(const kestrel_cohort* c, double* sweep_ms, uint64_t* sweeps, double* wall_ms);

/* The last fit's sweep-busy time (milliseconds with at least one sweep
 * running; the launches of the two slot groups overlap) and its number of
 * slot groups (option "streams"). */
int kestrel_last_fit_busy(const kestrel_cohort* c, double* busy_ms, uint32_t* groups);

/* The last kestrel_fit_pairs of a KESTREL_MIXED cohort: host milliseconds of
 * the FP64 condition pass (included in the fit's wall time) and the records
 * whose condition it estimated again (no counterpart in the reference). */
int kestrel_last_fit_recondition(const kestrel_cohort* c, double* ms, uint64_t* pairs);

/* The sweep form of the last kestrel_fit_pairs, which fixes its summation
 * order (records of different forms differ in the last bits): the site-table
 * sweep (one pair a lane; the mixed, soft-likelihood and tri/tetra-allelic
 * sweeps), the class-sorted sweep of 2-bit rows (FP64, the 2-allele part of a
 * tri/tetra-allelic cohort included; its other parts keep their sweeps), or a
 * pair's sites across a wavefront (option "pair_lanes").  No counterpart in
 * the reference (one summation order). */
enum { KESTREL_FORM_TABLE = 0, KESTREL_FORM_CLASS = 1, KESTREL_FORM_LANES = 2 };
int kestrel_last_fit_form(const kestrel_cohort* c, int* form);

/* Serving: the cohort stays device-resident and answers requests of query
 * samples (new 2-bit rows over the cohort's n_sites, fitted with the cohort's
 * allele frequencies and kept sites) against the whole cohort or a subset of
 * it.  The reference fits one batch per run and has no counterpart; each pair
 * is the fit of coefficients.rs:40-141 (start point 1/9, sqp.rs:49-93).
 *
 * Requests are admitted while others run (continuous batching), first come
 * first served, when a request slot, max_query_rows rows and max_pairs output
 * records (the device memory fixed at kestrel_serve_start) have room; the
 * admitted requests share the device pair by pair.  A pair's result does not
 * depend on what else is running: it is bit-identical alone and in any mix.
 *
 *   kestrel_serve_start     starts the engine thread (tune fixed per engine).
 *   kestrel_submit_queries  query_packed: n_query rows of row_stride_bytes;
 *                           cohort_subset: n_subset cohort sample indices, or
 *                           NULL for all samples.  out (caller-owned, kept
 *                           alive until the request is done) receives
 *                           n_query x n_subset records, out[q * n_subset + j]
 *                           with x = q (query index) and y = cohort index.
 *                           The query rows are copied; returns at once.
 *   kestrel_poll            *done = 1 once out is written; with latency_ms
 *                           (submission to completion).  A finished request is
 *                           forgotten after it was reported done once.
 *   kestrel_wait            blocks until the request is done.
 *   kestrel_serve_stop      fails pending and running requests, stops the
 *                           engine, frees its device memory (also done by
 *                           kestrel_cohort_destroy).
 * Submit, poll, wait and stats are internally synchronised and may come from
 * any thread; kestrel_serve_stop and kestrel_cohort_destroy must not overlap
 * another call on the cohort. */
typedef struct {
    uint64_t requests_done;
    uint64_t pairs_done;
    uint64_t sweeps;
    uint64_t busy_slot_sweeps; /* sum over sweeps of the slots holding a pair */
    uint64_t device_bytes;     /* device memory held by the engine */
    uint32_t slots;            /* pairs in flight at most */
    uint32_t live;             /* admitted, unfinished requests */
    uint32_t pending;          /* submitted, not yet admitted */
    uint32_t reserved;
} kestrel_serve_stats;

int kestrel_serve_start(kestrel_cohort* c, const kestrel_tune* tune, uint32_t max_query_rows, uint64_t max_pairs);
int kestrel_submit_queries(kestrel_cohort* c, const uint8_t* query_packed, uint32_t n_query, size_t row_stride_bytes,
                           const uint32_t* cohort_subset, size_t n_subset, kestrel_output* out, uint64_t* request_id);
int kestrel_poll(kestrel_cohort* c, uint64_t request_id, int* done, double* latency_ms);
int kestrel_wait(kestrel_cohort* c, uint64_t request_id, double* latency_ms);
int kestrel_serve_get_stats(kestrel_cohort* c, kestrel_serve_stats* stats);
int kestrel_serve_stop(kestrel_cohort* c);

void kestrel_cohort_destroy(kestrel_cohort* c);

/* Device arithmetic self-test (no counterpart in the reference): compares, on
 * n hashed inputs, the sweep's correctly rounded reciprocal with IEEE division
 * and its positive-domain log with the full log; *mismatches = differing bit
 * patterns (0 expected). */
int kestrel_selftest(uint64_t n, uint64_t seed, uint64_t* mismatches);

/* Achievable FP64 vector FMA throughput of this device in TFLOP/s (no
 * counterpart in the reference): a kernel of independent fma chains, used to
 * state the measured roofline next to the vendor peak. */
int kestrel_measure_fp64_peak(double* tflops);

/* The same for FP32 FMA chains as this library compiles FP32 code (no packed
 * FP32 instructions): the rate an FP32 form of the per-site terms could use
 * (no counterpart in the reference). */
int kestrel_measure_fp32_peak(double* tflops);

/* Timing of the per-site Hessian accumulation H = sum_s c'_s c'_s^T (the 45
 * terms of jacquard/grad_hess.rs:149-448) in one of the forms a sweep could
 * give it, on synthetic scaled rows c'_k = table[s % 256][k] * (table[s %
 * 256][9] * (1 + (pair % 64) / 64)) in FP32 (no counterpart in the reference:
 * it measures the matrix-core form against the vector forms).  form: 0 one
 * pair per lane in FP64, 1 one pair per lane in FP32, 2 one pair per
 * wavefront with its sites across the lanes and H by bf16 hi/lo matrix-core
 * products accumulated in FP32, 3 the same with a hi/mid/lo split, 4 form 1
 * and 5 form 2 without the Hessian (rows and gradient sums only), 6 form 2's
 * layout with the Hessian by FP32 FMA in each lane.  table:
 * 256 x 12 floats on the host; n_pairs a multiple of 256, n_sites of 4096
 * (the sites are summed in segments of 4096, as in the sweeps).  out (host):
 * per segment, n_pairs rows of 54 values (45 Hessian terms, lower triangle by
 * rows, then the 9 sums of c'), doubles for form 0 and floats otherwise;
 * *ms: the second of two launches. */
int kestrel_measure_hessian_form(int form, uint32_t n_pairs, uint32_t n_sites, const float* table, void* out,
                                 double* ms);

/* Achievable LDS read bandwidth of this device in TB/s (no counterpart in the
 * reference): conflict-free 16-byte reads by every lane, the bound of the CLS
 * per-pair kernel. */
int kestrel_measure_lds(double* tbs);

/* Plain device copy of `bytes` (two launches, the second timed): HBM
 * read+write GB/s (no counterpart in the reference; the reference kernel for
 * scaling hardware byte counters). */
int kestrel_measure_copy(uint64_t bytes, double* gbps);

const char* kestrel_last_error(void);
const char* kestrel_version(void);

#ifdef __cplusplus
}
#endif
#endif
