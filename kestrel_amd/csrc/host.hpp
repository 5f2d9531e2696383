// host.hpp -- host-side pieces shared by the C ABI sources: error
// reporting, device buffers, and the cohort handle.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
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
// Stops the serving engine of a cohort (if any) and frees its buffers.
void serve_release(ServeState* s);

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
    DevBuf<double> glrows;  // n_samples x v_pad x 3 likelihoods of kept sites (0 when missing)
    DevBuf<double> qmat;    // CLS Gram matrix (computed on first use)
    bool have_q = false;
    int64_t opt_slots = 0;
    int64_t opt_segment_sites = 4096;
    int64_t opt_serve_policy = 1;  // 0 round robin, 1 shortest remaining first
    int64_t opt_streams = 2;       // slot groups on separate streams in a fit
    hipStream_t stream2 = nullptr; // second stream of a fit
    double last_sweep_ms = 0, last_wall_ms = 0, last_busy_ms = 0;
    uint64_t last_sweeps = 0;
    int last_groups = 1;

    int set_segments() {
        // nseg = ceil(chunks / target); segments of equal length (the last
        // may be shorter) so that every block of a sweep has the same work
        const uint32_t chunks = v_pad / CHUNK;
        const uint32_t target = (uint32_t)std::max<int64_t>(CHUNK, opt_segment_sites) / CHUNK;
        nseg = std::max<uint32_t>(1, (chunks + target - 1) / target);
        const uint32_t per = (chunks + nseg - 1) / nseg;
        std::vector<uint32_t> b(nseg + 1);
        for (uint32_t k = 0; k <= nseg; k++) b[k] = std::min(chunks, k * per) * CHUNK;
        b[nseg] = v_pad;
        HIPCHK(seg.alloc(nseg + 1));
        HIPCHK(hipMemcpy(seg.p, b.data(), b.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        return KESTREL_OK;
    }

    ServeState* serve = nullptr;  // serving engine (kestrel_serve_start)

    ~kestrel_cohort() {
        serve_release(serve);
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
    }
};
