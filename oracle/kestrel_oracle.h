/*
 * kestrel_oracle.h -- CPU FP64 restatement of kestrel's pairwise Jacquard
 * maximum-likelihood path.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * compare against.  The product (kestrel_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the kestrel 0.1.2 source tree).  Two summation orders are provided for the
 * per-site reductions:
 *   KO_ORDER_REF : the reference's AVX-512 order -- 8 interleaved lanes over
 *                  the compacted (NaN-filtered) rows, 32-block tiles for
 *                  grad/Hess, scalar remainder, _mm512_reduce_add_pd tree.
 *   KO_ORDER_SEG : site-segmented sequential order -- each segment of sites is
 *                  summed sequentially (masked sites add exact zeros) and the
 *                  segment partials are folded left to right.  This is the
 *                  order the device kernels use, so device results can be
 *                  compared bit for bit.
 *   KO_ORDER_FAST: KO_ORDER_REF in the reference's AVX-512 loop shape (the
 *                  CPU baseline's timing form; same bits).
 *   KO_ORDER_LANES: the segments of KO_ORDER_SEG, each summed by 64 strided
 *                  lanes and a butterfly (the device's wavefront form).
 *
 * Pins (tests/test_oracle_*.py): the identity-state probabilities against an
 * exact-rational brute force, the two reference encodings of c_s against each
 * other, log against libm within 1 ulp, the QP against enumeration of all
 * supports, GMW81 against its reconstruction, objective / gradient / Hessian
 * against compensated sums and finite differences, allele frequencies against
 * the closed-form root, fits against their KKT conditions, CLS Q and brier
 * against numpy, and the golden fixture tests/golden/c1_small.npz.
 * Parity unpinned against the running Rust reference (no Rust toolchain here,
 * crates not vendored): exact iteration counts and last bits under the real
 * _mm512_reduce_add_pd code generation, active-set tie-breaks on degenerate
 * pairs, sites at exactly p = 1/(2N), and ryu's float text (restated as
 * shortest round-trip digits).
 */
#ifndef KESTREL_ORACLE_H
#define KESTREL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KO_ORDER_FAST: KO_ORDER_REF computed in the reference's own AVX-512 loop
 * shape (c_s rows built per 8-site block, objective with the getexp/getmant
 * log, Hessian in three 15-accumulator passes per tile; ko_lanes_fast.c):
 * the same bits as KO_ORDER_REF, for timing the CPU baseline.  Hard-call
 * cohorts on AVX-512 hosts; elsewhere it runs KO_ORDER_REF. */
/* KO_ORDER_LANES: the wavefront order of the device's one-pair-per-wavefront
 * sweep (option "pair_lanes" 64): per segment, 64 lanes each sum the sites
 * s0 + l + 64 t in order (a log per 16 steps and at the last), then a
 * butterfly adds the lanes (o = 32 .. 1: v[i] + v[i ^ o]); rows/valid/seg as
 * KO_ORDER_SEG. */
/* KO_ORDER_CLASS: the order of the device's class-sorted sweep (ko_class.c):
 * per segment, blocks of 8192 sites, each genotype-pair pattern's sites in
 * turn across 64 lanes, sums in the pattern's small basis, a butterfly over
 * the lanes, then the expansion into g and h; hard-call cohorts only. */
enum { KO_ORDER_REF = 0, KO_ORDER_SEG = 1, KO_ORDER_FAST = 2, KO_ORDER_LANES = 3, KO_ORDER_CLASS = 4 };

/* sqp.rs:16-47 Tuneables */
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
} ko_tune;

/* coefficients.rs:143-151 Output, plus the usable-site count n_xy */
typedef struct {
    uint32_t x, y;
    double jacquard[9];
    double obj;
    uint64_t iters;
    double cond;
    uint32_t n_sites;
    uint32_t pad;
} ko_output;

void ko_tune_new(ko_tune* t, uint64_t sqp_max_iter, uint64_t qp_max_iter);

/* log/scalar.rs:17-70 and log/avx512.rs:15-79 */
double ko_log(double d);
double ko_log_getmant(double d);

/* mixture.rs:96-236 (one site, arbitrary likelihood rows) */
void ko_mixture_row2(const double lx[3], const double ly[3], const double af[2], double out[9]);
/* iis.rs:94-104 + conditional.rs:10-50 : IIS mode (1..15) of genotype pair */
int ko_iis_mode(int i, int j, int k, int l);
/* iis.rs:106-277 accumulate_iis over conditional.rs IIS_2 (the generic path) */
void ko_mixture_generic2(const double lx[3], const double ly[3], const double af[2], double out[9]);
/* iis.rs:279-462 conditional_probabilities */
void ko_conditional_probabilities(double pi, double pj, double pk, double pl, int mode, double out[9]);

/* jacquard/objective.rs:8-52 ; rows are compacted (n x 9, row-major) */
double ko_jacq_obj_ref(const double* rows, size_t n, const double x[9], double eps);
/* jacquard/grad_hess.rs:12-449 */
void ko_jacq_gh_ref(const double* rows, size_t n, const double x[9], double eps, double g[9], double h[81]);

/* segmented order: rows are per site (V x 9), valid[s] != 0 marks usable sites,
 * seg has nseg+1 site boundaries. */
double ko_jacq_obj_seg(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg,
                       size_t nseg, const double x[9], double eps);
void ko_jacq_gh_seg(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                    const double x[9], double eps, double g[9], double h[81]);

/* wavefront order (KO_ORDER_LANES), arguments as the segmented order's */
double ko_jacq_obj_lanes(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg,
                         size_t nseg, const double x[9], double eps);
void ko_jacq_gh_lanes(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                      const double x[9], double eps, double g[9], double h[81]);

/* class order (ko_class.c): one pair's sums F = sum log L, G = sum c/L,
 * H = sum c c^T / L^2 (upper 45) and the usable-site count, from its codes
 * gx, gy (V each, 3 = missing), af (V x 2) and segment bounds; and the
 * expansion of a segment's 47 pattern sums into F, G, H. */
void ko_class_sums(const uint8_t* gx, const uint8_t* gy, const double* af, const uint32_t* seg, size_t nseg,
                   const double x[9], double eps, double* F, double G[9], double H[45], size_t* n);
void ko_class_expand(const double* R, double* F, double G[9], double H[45]);

/* cholesky.rs:6-65, 72-102, 105-140 (matrices are 9x9 row-major, first n used) */
void ko_modify_gmw(int n, const double* a, double* l, double* d);
void ko_unmodified(int n, const double* a, double* l, double* d);
void ko_solve_ldl(int n, const double* l, const double* d, double* b);

/* sqp.rs:111-250 ; N-variable (N <= 9), matrices N x N row-major */
uint64_t ko_solve_qp(int N, const double* q, const double* c, const double* y0, int modify, const ko_tune* t,
                     double* y);
/* sqp.rs:312-391 */
double ko_condition(int N, const double* x, const double* g, const double* h, double tol);

/* Full SQP of one pair (sqp.rs:49-93 with coefficients.rs:87-132).
 * order = KO_ORDER_REF : rows compacted n x 9
 * order = KO_ORDER_SEG : rows V x 9 with valid / seg */
void ko_fit_pair_rows(int order, const double* rows, size_t n_or_V, const uint8_t* valid, const uint32_t* seg,
                      size_t nseg, int self_pair, const ko_tune* t, ko_output* out);
/* The class order with tri/tetra-allelic parts: the 2-allele part's sums in
 * the class order (codes gx, gy and frequencies af2 of its V2 sites, one
 * segment), then the parts' rows (V x 9 with valid) per segment of pseg,
 * folded left to right after it. */
void ko_fit_pair_class_parts(const uint8_t* gx, const uint8_t* gy, const double* af2, uint32_t V2,
                             const double* prows, const uint8_t* pvalid, const uint32_t* pseg, size_t npseg,
                             int self_pair, const ko_tune* t, ko_output* out);
/* f, g, h (9 x 9) of one point in that order */
void ko_class_parts_eval(const uint8_t* gx, const uint8_t* gy, const double* af2, uint32_t V2, const double* prows,
                         const uint8_t* pvalid, const uint32_t* pseg, size_t npseg, const double x[9], double eps,
                         double* f, double g[9], double h[81]);

/* Build per-site c rows (V x 9) for a pair from 2-bit genotype codes (0,1,2 =
 * index of the unphased genotype, coefficients.rs:18-38; 3 = missing) using
 * mixture.rs:96-236 on one-hot likelihoods.  valid[s] = both called.
 * Returns n_xy. */
size_t ko_pair_rows_codes(const uint8_t* gx, const uint8_t* gy, const double* af, size_t V, double* rows,
                          uint8_t* valid);

/* allele/mod.rs:78-134 for biallelic sites from codes (one-hot GL rows, NaN
 * for missing).  codes is sample-major (N x V).  af_out is V x 2.
 * Returns 0, or -1 if some site has no called sample (lanevector.rs:73-76). */
int ko_allele_freqs_codes(const uint8_t* codes, uint32_t N, uint32_t V, uint64_t iterations, double maf,
                          int nthreads, double* af_out);
/* one site, arbitrary likelihood rows (n x 3, NaN rows skipped) */
int ko_allele_freq2_rows(const double* gl, size_t n, uint64_t iterations, double maf, double af[2],
                         uint64_t* iters);

/* Pair fits over a cohort of codes (N x V sample-major), af (V x 2).
 * pairs: npairs (x,y) entries, or NULL for the upper triangle incl. diagonal of
 * the first nsel samples, row-major (coefficients.rs:57,87).
 * seg/nseg describe site segments for KO_ORDER_SEG (ignored for REF).
 * Threads take pairs with an atomic counter. */
int ko_fit_pairs_codes(int order, const uint8_t* codes, uint32_t N, uint32_t V, const double* af,
                       const uint32_t* pairs, size_t npairs, uint32_t nsel, const uint32_t* seg, size_t nseg,
                       uint64_t iterations, int nthreads, ko_output* out);

/* Sites with A = 2, 3 or 4 alleles (G = A(A+1)/2 likelihoods in VCF order):
 * mixture.rs:8-60 over IIS_A, the allele-frequency fit of one site
 * (allele/objective.rs, grad_hess.rs lane2/3/4; QP cap A + 1; maf zeroing),
 * and the pair rows of one pair (lx, ly V x G, af V x A). */
void ko_mixture_generic(int A, const double* lx, const double* ly, const double* af, double out[9]);
int ko_allele_freqA_rows(int A, const double* gl, size_t n, uint64_t iterations, double maf, double* af,
                         uint64_t* iters);
size_t ko_pair_rows_generic(int A, const double* lx, const double* ly, const double* af, size_t V, double* rows,
                            uint8_t* valid);

/* Pair fits over re-bucketed 2-, 3- and 4-allele sites in the reference's
 * order (coefficients.rs:97-107), gl_A sample-major N x V_A x G_A, af_A
 * V_A x A; -1 if a pair names a sample >= N. */
int ko_fit_pairs_multi(uint32_t N, const double* gl2, const double* af2, uint32_t V2, const double* gl3,
                       const double* af3, uint32_t V3, const double* gl4, const double* af4, uint32_t V4,
                       const uint32_t* pairs, size_t npairs, uint64_t iterations, int nthreads, ko_output* out);

/* Soft likelihoods (vcf.rs:338-410 rows: 10^(gl - max) or 10^(-pl/10), NaN =
 * missing), sample-major N x V x 3.  Pair rows of one pair (V x 3 each), the
 * allele-frequency fit (allele/mod.rs:37-134) and the pair fits
 * (coefficients.rs:40-141) as their code counterparts. */
size_t ko_pair_rows_gl(const double* lx, const double* ly, const double* af, size_t V, double* rows,
                       uint8_t* valid);
int ko_allele_freqs_gl(const double* gl, uint32_t N, uint32_t V, uint64_t iterations, double maf, int nthreads,
                       double* af_out);
int ko_fit_pairs_gl(int order, const double* gl, uint32_t N, uint32_t V, const double* af, const uint32_t* pairs,
                    size_t npairs, uint32_t nsel, const uint32_t* seg, size_t nseg, uint64_t iterations, int nthreads,
                    ko_output* out);

/* CLS / Brier path of -f GT (coefficients.rs:166-320, cls.rs, quadratic.rs,
 * wrapper.rs:412-428).  codes N x V without missing calls. */
void ko_gt_allele_counts(const uint8_t* codes, uint32_t N, uint32_t V, double* af);
void ko_cls_m_matrices(const double* af, uint32_t V, double* m);
void ko_cls_gram(const double* m, uint32_t V, double* q);
int ko_cls_fit_pairs(const uint8_t* codes, uint32_t N, uint32_t V, const double* af, const uint32_t* pairs,
                     size_t npairs, uint32_t nsel, uint64_t iterations, int nthreads, ko_output* out, double* q_out);
/* Tri/tetra-allelic form: genotype indices in VCF order per bucket. */
void ko_gt_allele_counts_generic(int A, const uint8_t* idx, uint32_t N, uint32_t V, double* af);
void ko_cls_m_generic(int A, const double* af, uint32_t V, double* m);
void ko_cls_gram_multi(const double* m2, uint32_t V2, const double* m3, uint32_t V3, const double* m4, uint32_t V4,
                       double* q);
int ko_cls_fit_pairs_multi(uint32_t N, const uint8_t* idx2, uint32_t V2, const uint8_t* idx3, uint32_t V3,
                           const uint8_t* idx4, uint32_t V4, const double* af2, const double* af3, const double* af4,
                           const uint32_t* pairs, size_t npairs, uint32_t nsel, uint64_t iterations, int nthreads,
                           ko_output* out, double* q_out);

#ifdef __cplusplus
}
#endif
#endif
