// kestrel_internal.hpp -- shared device/host definitions of libkestrel_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kestrel_amd.h"
#include "sqp_device.hpp"

constexpr int CHUNK = 32;       // sites per LDS table chunk
constexpr int LOG_GROUP = 16;   // sites per product in the objective (one log per group)
constexpr int SWEEP_TPB = 256;  // threads (pairs) per sweep block
constexpr int NPART = 56;       // f, g[9], h[45], n per segment partial
constexpr uint32_t SITE_PAD = 256;  // rows padded to a multiple of this many sites
constexpr int CLASS_BLOCK = 8192;   // sites per block of the class-sorted sweep's lists
constexpr int CLASS_GROUP = 256;    // list entries per group (four chunks of 64)
constexpr int CLASS_STRIDE = CLASS_BLOCK + 5 * CLASS_GROUP;  // list entries per block (patterns group-aligned)
constexpr int CLASS_WAVES = 16;     // wavefronts (pairs) per class-sorted sweep workgroup (default)

// PH_WAIT: the slot holds a reservation (SlotState::pair) of a serving-queue
// entry that is not published yet
enum : uint32_t { PH_IDLE = 0, PH_INIT = 1, PH_TRIAL = 2, PH_RETIRED = 3, PH_WAIT = 4 };

// Per-slot SQP state between sweeps.
struct SlotState {
    int64_t pair;    // output record of the pair in flight, -1 when none
    uint32_t phase;  // PH_*
    uint32_t k;      // completed SQP iterations
    uint32_t bls;    // rejected trials of the current line search
    uint32_t req;    // request slot (serving queue)
    double x[9];     // current iterate
    double f;        // objective at x
    double p[9];     // search direction
    double t;        // sufficient-decrease slope c1 * g.p
    double alpha;    // current step
};

// Serving queue: a ring of pair entries (x row, y row, output record, request
// slot) published up to *tail; cap == 0 selects the batch pair list instead.
// Slots reserve entry indices with one atomic increment of the claim counter,
// which may run ahead of the tail by at most the number of slots; a slot whose
// index is not yet published waits (PH_WAIT) and reads it at a later step.
// After any step, every entry below min(claim counter, tail) has been read, so
// the host may overwrite ring positions up to that point.
struct PairQueue {
    const uint32_t* ent;
    uint64_t cap;
    const unsigned long long* tail;
    uint32_t* req_done;  // finished pairs per request slot
    int* busy;           // slots holding a pair
};

// Claims in runs (batch fits): the claim order is cut into runs of
// consecutive claims, each at least 256 long (the last may be shorter) and cut
// where the first sample changes, so that the 256 lanes of a sweep workgroup,
// which claim together from their workgroup's run, mostly fit pairs of one
// sample x.  runs[2r], runs[2r + 1]: claim indices [begin, end) of run r;
// wrun[2w], wrun[2w + 1]: the unclaimed part of workgroup w's run.
struct RunClaim {
    const uint64_t* runs = nullptr;
    uint64_t nruns = 0;
    unsigned long long* wrun = nullptr;
};

namespace klaunch {
hipError_t af(const uint8_t* packed, size_t row_bytes, uint32_t n_samples, uint32_t n_sites, const kdev::Tune& t,
              double maf, double* af_out, int* status, hipStream_t st);
// soft likelihoods: allele frequencies from site-major V x N x 3 rows, and
// the kept sites into sample-major rows (v_pad triples) + 2-bit mask rows
hipError_t af_gl(const double* gl, uint32_t n_samples, uint32_t n_sites, const kdev::Tune& t, double maf,
                 double* af_out, int* status, hipStream_t st, int alleles = 2);
// compact: 16-byte rows (the entries beside an exact 1.0, its position in the
// 2-bit code); *no_one is set when a called row has no entry equal to 1.
hipError_t gather_gl(const double* gl_site, uint32_t n_samples, const uint32_t* kept, uint32_t n_kept, double* rows,
                     uint32_t v_pad, uint32_t* mask, uint32_t row_words, int compact, int* no_one, hipStream_t st);
// tri/tetra-allelic parts: the sweep over a part of sites with 3 or 4 alleles
// (rows of A(A+1)/2 likelihoods, af of A per site), and the re-bucketing gather
hipError_t sweep_multi(int alleles, const uint32_t* geno, uint32_t row_words, const double* gl, uint32_t v_pad,
                       const double* table, const double* af, const uint32_t* seg_bounds, uint32_t nseg,
                       const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active, uint32_t nslots,
                       double eps, double* partial, hipStream_t st, const float* table32 = nullptr,
                       const uint32_t* codes = nullptr, const double* lut = nullptr);
// a hard-call part's rows of c_s per site and genotype pair (sweep_multi's
// lut with codes): part_lut_size doubles a site
size_t part_lut_size(int alleles);
hipError_t part_lut(int alleles, const double* table, uint32_t v_pad, double* lut, hipStream_t st);
// a part's 4-bit genotype indices (v_pad / 8 words a sample; 15 = missing)
// from its likelihood rows; *bad set when a called row is not one-hot
hipError_t part_codes(int alleles, const double* rows, const uint32_t* mask, uint32_t row_words, uint32_t v_pad,
                      uint32_t n_samples, uint32_t* codes, int* bad, hipStream_t st);
// dst[i] = (float)src[i] (the mixed form's FP32 coefficient tables; table32
// above selects that form of sweep_multi)
hipError_t round_f32(const double* src, float* dst, uint64_t n, hipStream_t st);
// *nonzero |= 1 where rows[0..n) holds a value other than zero
hipError_t any_nonzero(const double* rows, uint64_t n, int* nonzero, hipStream_t st);
// evaluation at points outside the sweeps' domain, the reference's arithmetic
// at every site (kind 0: 2-bit rows, 2 / 3: likelihood rows of that stride,
// 13 / 14: a 3- / 4-allele part); only slots with exact[slot] != 0
hipError_t sweep_exact(int kind, const uint32_t* geno, uint32_t row_words, const double* gl, uint32_t v_pad,
                       const double* af_or_table, const uint32_t* seg_bounds, uint32_t nseg, const uint32_t* sx,
                       const uint32_t* sy, const double* pt, const int* exact, uint32_t nslots, double eps,
                       double* partial, hipStream_t st);
// per-site accumulate_iis coefficients of a part (iis_table_size doubles a site)
int iis_table_size(int alleles);
hipError_t iis_table(int alleles, const double* af, uint32_t n_sites, double* table, hipStream_t st);
hipError_t gather_multi(const double* src2, const double* src3, const double* src4, uint32_t n_samples,
                        const uint32_t* map_site, const uint8_t* map_alleles, const uint8_t* gmap, uint32_t n_t,
                        int gp, double* rows, uint32_t v_pad, uint32_t* mask, uint32_t row_words, hipStream_t st,
                        int compact = 0, int* no_one = nullptr);
hipError_t gather(const uint8_t* src, size_t src_row_bytes, const uint32_t* kept, uint32_t n_kept,
                  uint32_t n_samples, uint32_t* dst, uint32_t dst_row_words, hipStream_t st,
                  const uint32_t* dst_rows = nullptr);
hipError_t sweep(const uint32_t* geno, uint32_t row_words, const double* af, const uint32_t* seg_bounds,
                 uint32_t nseg, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active,
                 uint32_t nslots, double eps, double* partial, hipStream_t st,
                 const double* gl = nullptr, uint32_t v_pad = 0, int gl_stride = 3, bool skip_masked = false,
                 bool mixed = false);
// FP64 fits of 2-bit or biallelic likelihood-row cohorts (gl != nullptr), one
// pair per wavefront with its sites across the lanes (option "pair_lanes"
// 64): partials in sweep's layout, another summation order
hipError_t sweep_lanes(const uint32_t* geno, uint32_t row_words, const double* af, const uint32_t* seg_bounds,
                       uint32_t nseg, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active,
                       uint32_t nslots, double eps, double* partial, hipStream_t st, const double* gl = nullptr,
                       uint32_t v_pad = 0, int gl_stride = 3);
// class-sorted sweep of 2-bit FP64 fits: per slot the pattern lists of its
// pair (built where built[slot] differs from (sx, sy); nblk blocks of
// CLASS_STRIDE entries, counts 8 uint16 a block), then the sweep over segments
// whose bounds are multiples of CLASS_BLOCK; partials in sweep's layout.
// Both cover the slots slot0 .. slot0 + nchunk - 1, whose lists (and built
// marks) sit at 0 .. nchunk - 1 of the arrays; the other arrays are indexed
// by slot (points and partials with stride nslots)
hipError_t class_build(const uint32_t* geno, uint32_t row_words, const uint32_t* sx, const uint32_t* sy,
                       const int* active, uint32_t nslots, uint32_t slot0, uint32_t nchunk, uint32_t nblk,
                       uint16_t* lists, uint16_t* counts, uint2* built, hipStream_t st);
hipError_t sweep_class(const double* af, uint32_t v_pad, const uint32_t* seg_bounds, uint32_t nseg,
                       const uint16_t* lists, const uint16_t* counts, uint32_t nblk, const double* pt,
                       const int* active, uint32_t nslots, uint32_t slot0, uint32_t nchunk, double eps,
                       double* partial, hipStream_t st, int waves = CLASS_WAVES);
hipError_t step(SlotState* stt, uint32_t* sx, uint32_t* sy, double* pt, int* active, uint32_t nslots,
                double* partial, uint32_t nseg, const uint32_t* pairs, uint64_t npairs, const uint64_t* order,
                unsigned long long* next_pair, kestrel_output* out, const kdev::Tune& t, int* n_live,
                const PairQueue& q, int* live_out, hipStream_t st, const RunClaim& rc = RunClaim{});
// code-3 calls of the first n_sites sites of every row (missing-call count)
hipError_t count_missing(const uint32_t* geno, uint32_t n_samples, uint32_t row_words, uint32_t n_sites,
                         unsigned long long* count, hipStream_t st);
hipError_t selftest(uint64_t n, uint64_t seed, unsigned long long* bad, hipStream_t st);
hipError_t calib_fp64(uint32_t blocks, uint32_t iters, double* sink, hipStream_t st);
hipError_t calib_fp32(uint32_t blocks, uint32_t iters, float* sink, hipStream_t st);
hipError_t calib_lds(uint32_t blocks, uint32_t iters, double* sink, hipStream_t st);
hipError_t copy(const void* src, void* dst, uint64_t n16, hipStream_t st);
hipError_t complete_sites(const uint8_t* packed, size_t row_bytes, uint32_t n_samples, uint32_t n_sites, uint8_t* flag,
                          hipStream_t st);
hipError_t gt_count(const uint32_t* geno, uint32_t row_words, uint32_t n_samples, uint32_t n_sites, double* af,
                    hipStream_t st);
// CLS Gram matrix over the 2-allele sites' m rows (from af2), then the 3- and
// 4-allele sites' precomputed tables m3 / m4; z holds gram_tiles(...) x 45 x 8.
hipError_t gram(const double* af2, uint32_t n2, const double* m3, uint32_t n3, const double* m4, uint32_t n4,
                double* z, double* q, hipStream_t st);
uint64_t gram_tiles(uint32_t n2, uint32_t n3, uint32_t n4);
hipError_t gt_count_multi(int alleles, const uint32_t* geno, uint32_t row_words, uint32_t n_samples, uint32_t n_sites,
                          double* af, hipStream_t st);
hipError_t cls_m(int alleles, const double* af, uint32_t n_sites, double* m, hipStream_t st);
hipError_t cls_multi(int alleles, const uint32_t* geno, uint32_t row_words, const double* mtab, uint32_t n_sites,
                     const uint32_t* pairs, uint64_t npairs, int add_prev, double* out, hipStream_t st);
// csum: npairs x 9 doubles of scratch (the 2-allele sums)
hipError_t cls_pairs(const uint32_t* geno, uint32_t row_words, const double* af, uint32_t n_sites, const double* c34,
                     uint32_t v_total, const double* q, const uint32_t* pairs, uint64_t npairs, const kdev::Tune& t,
                     double* csum, kestrel_output* out, hipStream_t st);
hipError_t finalize(const double* partial, uint32_t nseg, uint32_t nslots, uint32_t npts, double* f, double* g,
                    double* h, uint32_t* nsites, hipStream_t st);
// condition estimates at the points pt (SoA, nslots apart) from FP64 partials
hipError_t recondition(const double* partial, uint32_t nseg, uint32_t nslots, uint32_t npts, const double* pt,
                       double tol, double* cond, hipStream_t st);
}  // namespace klaunch
