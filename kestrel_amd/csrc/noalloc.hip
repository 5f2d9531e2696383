// noalloc.hip -- kestrel_evaluate_noalloc: the objective / gradient / Hessian
// evaluation of kestrel_evaluate for callers that cannot allocate device
// memory.  Inputs and outputs are the caller's device buffers; the sweep's
// slot arrays, segment partials and (class-sorted sweep) pattern lists live in
// `scratch`, a fixed array in device memory that each GPU holds until the
// process ends.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host.hpp"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;  // the stated maximum
constexpr size_t SCRATCH_HEAD = 256;                // error counter, then the slot arrays
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// Per-device address of `scratch`, looked up once per device with that
// device current (each GPU holds its own copy of the array).
std::mutex g_mu;
unsigned char* g_ptr[MAX_DEVICES] = {};

int scratch_ptr(int device, unsigned char** out) {
    if (device < 0 || device >= MAX_DEVICES) return fail(KESTREL_EDEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != device) HIPCHK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) HIPCHK(hipSetDevice(cur));
        HIPCHK(e);
        g_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_ptr[device];
    return KESTREL_OK;
}

// Layout of one call in scratch (after the 256-byte head): sx, sy (ns u32),
// active, exact (ns int), points (9 x ns doubles, SoA), partials
// (nseg x 56 x ns); with the class-sorted sweep then the lists' built marks
// (ns x 8 bytes), counts (ns x nblk x 16 bytes) and lists (ns x nblk x
// CLASS_STRIDE x 2 bytes).
size_t slot_bytes(uint32_t nseg, uint32_t nblk) {
    return 4 * sizeof(uint32_t) + 9 * sizeof(double) + (size_t)nseg * NPART * 8 +
           (nblk ? 8 + (size_t)nblk * (16 + 2 * CLASS_STRIDE) : 0);
}

__global__ __launch_bounds__(64) void clear_head_kernel(unsigned int* __restrict__ err) {
    if (threadIdx.x == 0) *err = 0u;
}

// scratch words set to a value (a kernel rather than a memset, so that the
// call can be captured into a graph)
__global__ __launch_bounds__(256) void fill_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

// The caller's pairs (n x 2) and points (n x 9) into the sweep's slot arrays;
// slots past n are inactive.  A point inside the sweeps' domain goes to the
// sweeps (active), one outside it to the exact kernel (exact).  A pair naming
// a sample beyond the cohort is counted in *err and left to neither.
__global__ __launch_bounds__(256) void pack_kernel(const uint32_t* __restrict__ pairs,
                                                   const double* __restrict__ points, uint32_t n, uint32_t ns,
                                                   uint32_t n_samples, double eps, double lmax,
                                                   uint32_t* __restrict__ sx, uint32_t* __restrict__ sy,
                                                   int* __restrict__ active, int* __restrict__ exact,
                                                   double* __restrict__ pt, unsigned int* __restrict__ err) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ns) return;
    uint32_t x = 0, y = 0;
    int act = 0, ex = 0;
    double d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = i < n ? points[(size_t)i * 9 + k] : 0.0;
    if (i < n) {
        x = pairs[2 * (size_t)i];
        y = pairs[2 * (size_t)i + 1];
        if (x < n_samples && y < n_samples) {
            act = point_in_domain(d, eps, lmax) ? 1 : 0;
            ex = 1 - act;
        } else {
            atomicAdd(err, 1u);
            x = 0;
            y = 0;
        }
    }
    sx[i] = x;
    sy[i] = y;
    active[i] = act;
    exact[i] = ex;
#pragma unroll
    for (int k = 0; k < 9; k++) pt[(size_t)k * ns + i] = d[k];
}

}  // namespace

extern "C" int kestrel_evaluate_noalloc(kestrel_cohort* c, const uint32_t* pairs, const double* points, size_t n,
                                        double eps, double* f, double* g, double* h, uint32_t* n_sites) {
    if (!c || !pairs || !points || !f || !g || !h || !n_sites) return fail(KESTREL_EINVAL, "null argument");
    if (c->gt_parts[0] || c->gt_parts[1])
        return fail(KESTREL_EINVAL, "a -f GT cohort with 3- or 4-allele sites has no likelihood evaluation");
    if (n == 0) return KESTREL_OK;
    const bool cls = c->class_form();  // points inside the domain by the class-sorted sweep, as kestrel_evaluate
    // partial segments: the exact kernel's layout (the cohort's segments) or
    // the class-sorted form's (its one segment, then the parts'), whichever is longer
    const uint32_t nseg = std::max(c->nseg_total(), cls ? c->nseg_class_total() : 0u);
    const uint32_t nblk = cls ? c->class_blocks() : 0;
    const uint64_t ns = ((uint64_t)n + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB;
    const uint64_t cap = (SCRATCH_BYTES - SCRATCH_HEAD) / slot_bytes(nseg, nblk) / SWEEP_TPB * SWEEP_TPB;
    if (ns > cap)
        return fail(KESTREL_EINVAL,
                    "kestrel_evaluate_noalloc: %zu points exceed its %llu-point limit for this cohort (%u segments, "
                    "%u list blocks, %zu MiB of scratch)",
                    n, (unsigned long long)cap, nseg, nblk, SCRATCH_BYTES >> 20);
    unsigned char* base = nullptr;
    int rc = scratch_ptr(c->device, &base);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    unsigned int* err = reinterpret_cast<unsigned int*>(base);
    uint32_t* sx = reinterpret_cast<uint32_t*>(base + SCRATCH_HEAD);
    uint32_t* sy = sx + ns;
    int* act = reinterpret_cast<int*>(sy + ns);
    int* ex = act + ns;
    // 8-byte alignment for the doubles: 4 ns words, ns a multiple of 256
    double* pt = reinterpret_cast<double*>(ex + ns);
    double* part = pt + 9 * ns;
    // class lists: 8-byte built marks, then 16-byte counts (8-byte aligned
    // after the partials; ns a multiple of 256 keeps 16-byte alignment)
    uint2* built = reinterpret_cast<uint2*>(part + (size_t)nseg * NPART * ns);
    uint16_t* counts = reinterpret_cast<uint16_t*>(built + ns);
    uint16_t* lists = counts + ns * nblk * 8;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(clear_head_kernel, dim3(1), dim3(64), 0, st, err);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)(ns / 256)), dim3(256), 0, st, pairs, points, (uint32_t)n,
                       (uint32_t)ns, c->n_samples, eps, c->lmax, sx, sy, act, ex, pt, err);
    HIPCHK(hipGetLastError());
    if (cls) {
        // every slot's lists are built anew; the sweep writes its one segment,
        // the exact kernel the cohort's: the others start at zero
        hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(256), 0, st, reinterpret_cast<uint32_t*>(built), 2 * ns,
                           0xffffffffu);
        hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, st, reinterpret_cast<uint32_t*>(part),
                           (size_t)nseg * NPART * ns * 2, 0u);
        HIPCHK(hipGetLastError());
        HIPCHK(klaunch::class_build(c->geno.p, c->row_words, sx, sy, act, (uint32_t)ns, 0, (uint32_t)ns, nblk, lists,
                                    counts, built, st));
        HIPCHK(klaunch::sweep_class(c->af.p, c->v_pad, c->seg_cls.p, c->nseg_cls, lists, counts, nblk, pt, act,
                                    (uint32_t)ns, 0, (uint32_t)ns, eps, part, st, (int)c->opt_class_waves));
    } else {
        HIPCHK(klaunch::sweep(c->geno.p, c->row_words, c->af.p, c->seg.p, c->nseg, sx, sy, pt, act, (uint32_t)ns,
                              eps, part, st, c->soft ? c->glrows.p : nullptr, c->v_pad, c->gl_stride, false,
                              c->mixed));
    }
    size_t off = cls ? c->nseg_cls : c->nseg;  // the tri/tetra-allelic parts' first partial segment
    for (const auto& p : c->parts) {
        HIPCHK(klaunch::sweep_multi(p->alleles, p->mask.p, p->row_words, p->rows.p, p->v_pad, p->table.p, p->af.p,
                                    p->seg.p, p->nseg, sx, sy, pt, act, (uint32_t)ns, eps, part + off * NPART * ns, st,
                                    c->mixed ? p->table32.p : nullptr, p->codes.n ? p->codes.p : nullptr, p->lut.p));
        off += p->nseg;
    }
    // points outside the sweeps' domain (the kernel leaves every other slot)
    rc = launch_exact(c, sx, sy, pt, ex, (uint32_t)ns, eps, part, st);
    if (rc) return rc;
    HIPCHK(klaunch::finalize(part, nseg, (uint32_t)ns, (uint32_t)n, f, g, h, n_sites, st));
    unsigned int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, err, sizeof bad, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (bad) return fail(KESTREL_EINVAL, "%u pairs name a sample beyond the cohort's %u", bad, c->n_samples);
    return KESTREL_OK;
}
