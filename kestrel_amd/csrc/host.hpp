// host.hpp -- host-side pieces shared by the C ABI sources: error
// reporting, device buffers, and the cohort handle.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "kestrel_internal.hpp"

// Sets the calling thread's error message (kestrel_last_error) and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// The calling thread's current error message.
const std::string& last_error_message();

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? KESTREL_ENOMEM : KESTREL_EDEVICE, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    }
};

kdev::Tune to_dev(const kestrel_tune& t);

struct ServeState;

// Segment bounds of a part of v_pad sites: nseg = ceil(chunks / target),
// segments of equal length (the last may be shorter) so that every block of a
// sweep has the same work.
inline std::vector<uint32_t> segment_bounds(uint32_t v_pad, int64_t segment_sites) {
    const uint32_t chunks = v_pad / CHUNK;
    const uint32_t target = (uint32_t)std::max<int64_t>(CHUNK, segment_sites) / CHUNK;
    const uint32_t nseg = std::max<uint32_t>(1, (chunks + target - 1) / target);
    const uint32_t per = (chunks + nseg - 1) / nseg;
    std::vector<uint32_t> b(nseg + 1);
    for (uint32_t k = 0; k <= nseg; k++) b[k] = std::min(chunks, k * per) * CHUNK;
    b[nseg] = v_pad;
    return b;
}

// Segment size of a tri/tetra-allelic part: as many segments as the main part
// (main_nseg) so that the part's sweep launches fill the device like the main
// one, rounded up to whole chunks, at most the cohort's segment size.
inline int64_t part_segment_sites(uint32_t part_v_pad, uint32_t main_nseg, int64_t segment_sites) {
    const uint64_t per = ((uint64_t)part_v_pad + main_nseg - 1) / std::max<uint32_t>(1, main_nseg);
    const int64_t rounded = (int64_t)((per + CHUNK - 1) / CHUNK * CHUNK);
    return std::min<int64_t>(segment_sites, std::max<int64_t>(CHUNK, rounded));
}

// Sites with 3 or 4 alleles after zero-frequency removal (likelihood rows of
// A(A+1)/2 genotypes); their segments follow the main part's in a sweep.
struct SitePart {
    int alleles = 0;
    uint32_t n_sites = 0, v_pad = 0, row_words = 0, nseg = 0;
    DevBuf<double> rows;    // n_samples x v_pad x G
    DevBuf<uint32_t> mask;  // n_samples x row_words 2-bit rows, 3 = missing
    DevBuf<double> af;      // v_pad x alleles
    DevBuf<double> table;   // v_pad x the mode coefficients of every genotype pair (A >= 3)
    DevBuf<float> table32;  // the same rounded to FP32 (precision KESTREL_MIXED)
    DevBuf<uint32_t> seg;
    DevBuf<uint32_t> codes;  // hard calls (every called row one-hot): 4-bit genotype indices, else empty
    DevBuf<double> lut;      // with codes: each site's c_s rows of every genotype pair (part_lut)
};
// The 3- or 4-allele sites of a -f GT cohort (kestrel_cohort_create_gt_multi):
// genotype indices in 4-bit rows (8 sites per word, sample-major, padding 15),
// frequencies and the sites' m tables (n_sites x G x G x 9).
struct GtPart {
    int alleles = 0;
    uint32_t n_sites = 0, row_words = 0;
    DevBuf<uint32_t> geno;
    DevBuf<double> af;
    DevBuf<double> mtab;
};
// option "pair_lanes" -1: fits of at most this many pairs take the
// one-pair-per-wavefront sweep
constexpr size_t LANES_AUTO_PAIRS = 4096;
// The mixed-precision form's condition estimates: a converged record whose
// estimate from the FP32-summed Hessian is not below this is estimated again
// from an FP64 evaluation at its final point.
constexpr double RECOND_MIN = 100.0;
// ... and so is one whose strict-complementarity decision is marginal (the
// step reports -condition): within this distance of the test's threshold.
constexpr double RECOND_MARGIN = 1e-6;

struct kestrel_cohort;
// Recomputes the condition of the records that need it (RECOND_MIN) from FP64
// sweeps over geno (row_words per row, segments seg[0..nseg]); rows(i, x, y)
// gives the geno rows of record i.  Ordered on st, blocks until done.
int recondition_records(kestrel_cohort* c, const uint32_t* geno, uint32_t row_words, const uint32_t* seg,
                        uint32_t nseg, const kestrel_tune& tune, kestrel_output* recs, size_t n,
                        const std::function<void(size_t, uint32_t&, uint32_t&)>& rows, hipStream_t st,
                        uint64_t* n_done = nullptr);

// The sweeps' domain: eps in [1e-18, 1e14] and a point of finite,
// non-negative coefficients with eps + lmax^2 sum(D) <= 1e15.  There every
// L = eps + c.D (c >= 0, c_k <= lmax^2) lies in [1e-18, 1e15], so 1/L and
// the product of 16 L are positive normal numbers; points outside it are
// evaluated by the exact kernel.
__host__ __device__ inline bool eps_in_domain(double eps) { return eps >= 1e-18 && eps <= 1e14; }
__host__ __device__ inline bool point_in_domain(const double* d, double eps, double lmax) {
    if (!eps_in_domain(eps)) return false;
    double s = 0.0;
    for (int k = 0; k < 9; k++) {
        if (!(d[k] >= 0.0 && d[k] <= 1e15)) return false;
        s += d[k];
    }
    return eps + lmax * lmax * s <= 1e15;
}

// Exact-arithmetic evaluation of the slots with exact[slot] != 0 over every
// part of the cohort into the sweeps' partial layout.
int launch_exact(kestrel_cohort* c, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* exact,
                 uint32_t ns, double eps, double* partial, hipStream_t st);

// Stops the serving engine of a cohort (if any) and frees its buffers.
void serve_release(ServeState* s);

// Per-slot pattern lists of the class-sorted sweep (class_sweep.hip), for
// cap slots (launch_sweeps sweeps more slots than that in chunks of cap).
struct ClassLists {
    DevBuf<uint16_t> lists, counts;
    DevBuf<uint2> built;
    uint32_t nblk = 0, cap = 0;
    // device bytes of one slot's lists, counts and built mark
    static uint64_t slot_bytes(uint32_t nblocks) { return (uint64_t)nblocks * (2 * CLASS_STRIDE + 16) + 8; }
    int alloc(uint32_t nslots, uint32_t nblocks, hipStream_t st) {
        nblk = nblocks;
        cap = nslots;
        HIPCHK(lists.alloc((size_t)nslots * nblk * CLASS_STRIDE));
        HIPCHK(counts.alloc((size_t)nslots * nblk * 8));
        HIPCHK(built.alloc(nslots));
        // no pair's lists are built yet (0xffffffff names no sample)
        HIPCHK(hipMemsetAsync(built.p, 0xff, (size_t)nslots * sizeof(uint2), st));
        return 0;
    }
};

// Slots whose class lists fit the device memory that one fit, evaluation or
// serving engine gives them (limit_mib MiB if set, else a quarter of the free
// memory, at most 16 GiB; shared by `groups` slot groups), a multiple of 256;
// 0 when 256 do not fit.
inline uint64_t class_slot_cap(uint32_t nblocks, int groups, int64_t limit_mib) {
    uint64_t budget = (uint64_t)limit_mib << 20;
    if (limit_mib <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
        budget = std::min<uint64_t>(free_b / 4, 16ull << 30);
    }
    return budget / (uint64_t)std::max(1, groups) / ClassLists::slot_bytes(nblocks) / 256 * 256;
}

struct kestrel_cohort {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    uint32_t n_samples = 0, n_sites = 0, n_kept = 0, v_pad = 0, row_words = 0;
    DevBuf<uint32_t> geno;  // n_samples x row_words packed kept sites
    DevBuf<double> af;      // v_pad x 2 kept (p, q), padding (0.5, 0.5)
    DevBuf<uint32_t> seg;   // segment bounds
    uint32_t nseg = 0;
    std::vector<double> af_all;
    std::vector<uint8_t> kept_mask;
    bool gt = false;        // created by kestrel_cohort_create_gt (CLS path)
    bool soft = false;      // created by kestrel_cohort_create_gl: likelihood rows in glrows
    DevBuf<double> glrows;  // n_samples x v_pad x gl_stride likelihoods of kept sites (0 when missing)
    int gl_stride = 3;      // 3: the three likelihoods; 2: the two beside an exact 1.0 (its position in geno)
    DevBuf<double> qmat;    // CLS Gram matrix (computed on first use)
    bool have_q = false;
    int64_t opt_slots = 0;
    int64_t opt_segment_sites = 4096;
    int64_t opt_serve_policy = 1;  // 0 round robin, 1 shortest remaining first
    int64_t opt_streams = 2;       // slot groups on separate streams in a fit
    int64_t opt_batch = 8;         // sweep/step iterations queued per host synchronisation
    int64_t opt_graph = 0;         // 1: the batch is captured once into a HIP graph per fit and relaunched
    int64_t opt_skip_masked = -1;  // claims in runs + sweeps passing over sites where a wavefront's shared
                                   // sample is missing: 1 on, 0 off, -1 when missing_frac >= 0.02
    double missing_frac = 0.0;     // missing calls among the kept sites (2-bit cohorts)
    double lmax = 1.0;             // largest likelihood of the cohort's rows (1 for hard calls)
    bool mixed = false;            // precision KESTREL_MIXED: sweeps by sweep_mixed
    int64_t opt_pair_tile = -1;    // claim order in tiles of this many samples (0 list order, -1 automatic)
    int64_t opt_pair_lanes = 1;    // lanes a pair: 1, 64 (its sites across a wavefront; FP64 fits), -1 auto
    int64_t opt_class_sweep = 1;   // FP64 fits of 2-bit cohorts by the class-sorted sweep (0: the table sweep)
    int64_t opt_class_waves = CLASS_WAVES;  // its wavefronts (pairs) per workgroup: 8, 12, 14 or 16
    int64_t opt_class_list_mib = 0;  // device memory its lists may take (0: automatic)
    DevBuf<uint32_t> seg_cls;      // its segment bounds (multiples of CLASS_BLOCK; one segment)
    uint32_t nseg_cls = 0;
    hipStream_t stream2 = nullptr; // second stream of a fit
    double last_sweep_ms = 0, last_wall_ms = 0, last_busy_ms = 0;
    uint64_t last_sweeps = 0;
    int last_groups = 1;
    uint32_t last_slots = 0;                        // pairs in flight per group in the last fit
    int last_form = KESTREL_FORM_TABLE;             // the last fit's sweep form
    double last_exhaust_ms = 0, last_end_ms = 0;    // device ms until the first slot went idle / the last sweep
    double last_recond_ms = 0;                      // mixed form: host ms of the FP64 condition pass
    uint64_t last_recond_pairs = 0;                 // mixed form: records whose condition it recomputed

    // tri/tetra-allelic cohorts (kestrel_cohort_create_gl_multi)
    std::vector<std::unique_ptr<SitePart>> parts;
    bool multi = false;
    uint32_t part_sites[3] = {0, 0, 0};   // sites with 2, 3, 4 alleles after zero-frequency removal
    std::vector<double> af_in[3];         // frequencies of the input buckets (2, 3, 4 alleles)
    std::unique_ptr<GtPart> gt_parts[2];  // -f GT cohorts: the 3- and 4-allele sites (null when none)

    uint32_t nseg_total() const {
        uint32_t n = nseg;
        for (const auto& p : parts) n += p->nseg;
        return n;
    }
    // partial segments of the class-sorted form: its one segment of the
    // 2-allele sites, then those of the tri/tetra-allelic parts
    uint32_t nseg_class_total() const {
        uint32_t n = nseg_cls;
        for (const auto& p : parts) n += p->nseg;
        return n;
    }

    int set_segments() {
        std::vector<uint32_t> b = segment_bounds(v_pad, opt_segment_sites);
        nseg = (uint32_t)b.size() - 1;
        HIPCHK(seg.alloc(nseg + 1));
        HIPCHK(hipMemcpy(seg.p, b.data(), b.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        const uint32_t bc[2] = {0, v_pad};
        nseg_cls = 1;
        HIPCHK(seg_cls.alloc(2));
        HIPCHK(hipMemcpy(seg_cls.p, bc, sizeof bc, hipMemcpyHostToDevice));
        for (auto& p : parts) {
            std::vector<uint32_t> pb = segment_bounds(p->v_pad, part_segment_sites(p->v_pad, nseg, opt_segment_sites));
            p->nseg = (uint32_t)pb.size() - 1;
            HIPCHK(p->seg.alloc(p->nseg + 1));
            HIPCHK(hipMemcpy(p->seg.p, pb.data(), pb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        return KESTREL_OK;
    }

    ServeState* serve = nullptr;  // serving engine (kestrel_serve_start)

    // the class-sorted sweep covers FP64 evaluations of 2-bit cohorts
    // FP64 fits of 2-bit rows (the 2-allele part of a tri/tetra-allelic
    // cohort too, whose other parts keep their sweeps)
    bool class_form() const { return opt_class_sweep == 1 && !soft && !mixed && !gt; }
    uint32_t class_blocks() const { return (v_pad + CLASS_BLOCK - 1) / CLASS_BLOCK; }

    ~kestrel_cohort() {
        serve_release(serve);
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
    }
};
