// kernels.hip -- gfx950 kernels of the pairwise Jacquard maximum-likelihood fit.
//
//   af_kernel      per-site allele-frequency fit (allele/mod.rs:78-134), one
//                  site per thread, 8-lane summation order of the reference.
//   gather_kernel  repack the kept (segregating) sites of the 2-bit cohort.
//   sweep_fp64     one evaluation pass: for 64 pairs per wavefront and one
//                  segment of sites, the objective sum of log(eps + c_s.D)
//                  (jacquard/objective.rs:42-72), the gradient sum of c_s / L
//                  and the 45 Hessian sums of (c_s/L)(c_s/L)^T
//                  (jacquard/grad_hess.rs:89-449), fused in one sweep.  c_s is
//                  read from an LDS table built per 32-site chunk with the
//                  exact expressions of mixture.rs:96-236.
//   step_kernel    per-slot SQP state machine (sqp.rs:49-93, 270-310): folds
//                  the segment partials, accepts or backtracks, runs the 9x9
//                  active-set QP, writes finished pairs and claims the next
//                  pair from an atomic counter (work stealing).
//   finalize_eval  folds partials into (f, g, H) for explicit evaluation points.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "kestrel_internal.hpp"
#include "sqp_device.hpp"

using namespace kdev;

namespace {

// ------------------------------------------------------------ allele fit

__device__ __forceinline__ double radd8(const double* v) {
    // Rust stdarch _mm512_reduce_add_pd: 256-bit halves, 128-bit halves, pair.
    const double a0 = v[0] + v[4], a1 = v[1] + v[5], a2 = v[2] + v[6], a3 = v[3] + v[7];
    const double b0 = a0 + a2, b1 = a1 + a3;
    return b0 + b1;
}

// Likelihood rows of one site: codes (one-hot rows of the 2-bit calls) or
// linear-scale likelihoods (site-major, samples consecutive); load() returns
// false for a row that lanevector.rs:51-53 skips (a missing call / a NaN).
struct CodeRows {
    const uint8_t* col;  // byte of this site in sample row 0
    size_t stride;       // row stride in bytes
    uint32_t shift;      // bit offset of the site inside its byte
    __device__ __forceinline__ bool load(uint32_t i, double* l) const {
        const uint32_t c = (col[(size_t)i * stride] >> shift) & 3u;
        l[0] = c == 0 ? 1.0 : 0.0;
        l[1] = c == 1 ? 1.0 : 0.0;
        l[2] = c == 2 ? 1.0 : 0.0;
        return c != 3;
    }
};

// Likelihood rows of a site with G genotypes, samples consecutive.
template <int G>
struct GlRows {
    const double* base;  // likelihoods of sample 0 at this site
    __device__ __forceinline__ bool load(uint32_t i, double* l) const {
        bool nan = false;
#pragma unroll
        for (int k = 0; k < G; k++) {
            l[k] = base[(size_t)G * i + k];
            nan |= isnan(l[k]);
        }
        return !nan;
    }
};

// allele/objective.rs and allele/grad_hess.rs for A alleles (the lane2/3/4
// functions are one pattern): G = A(A+1)/2 likelihoods per row in VCF order,
// genotype index j(j+1)/2 + i for alleles i <= j.
template <int A>
struct Genotypes {
    static constexpr int G = A * (A + 1) / 2;
    int gi[G], gj[G];
    constexpr Genotypes() : gi(), gj() {
        int g = 0;
        for (int j = 0; j < A; j++)
            for (int i = 0; i <= j; i++, g++) {
                gi[g] = i;
                gj[g] = j;
            }
    }
};

__host__ __device__ constexpr int gl_index(int a, int b) { return a < b ? b * (b + 1) / 2 + a : a * (a + 1) / 2 + b; }

template <class Rows, int A>
struct AlleleSite {
    static constexpr int G = A * (A + 1) / 2;
    Rows rows;
    uint32_t n_samples;
    double eps;

    // Visits the usable rows as LaneVectorBuilder lays them out
    // (lanevector.rs:99-139): full blocks of 8 in order, lane j taking the
    // j-th row of a block, then the remainder rows.
    template <class Lane, class Rem>
    __device__ __forceinline__ uint32_t visit(Lane lane, Rem rem) const {
        uint32_t idx[8];
        uint32_t cnt = 0, nb = 0;
        for (uint32_t i = 0; i < n_samples; i++) {
            double l[G];
            if (!rows.load(i, l)) continue;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if ((uint32_t)j == cnt) idx[j] = i;
            if (++cnt == 8) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    rows.load(idx[j], l);
                    lane(j, l);
                }
                cnt = 0;
                nb++;
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((uint32_t)j < cnt) {
                double l[G];
                rows.load(idx[j], l);
                rem(l);
            }
        return nb * 8 + cnt;
    }

    // allele/objective.rs:85-209: p = eps + sum_g w_g l_g, w = x_i x_i or
    // 2 x_i x_j, in genotype order
    __device__ double obj(const double* x) const {
        constexpr Genotypes<A> gt;
        double w[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (gt.gi[g] == gt.gj[g]) {
                w[g] = x[gt.gi[g]] * x[gt.gi[g]];
            } else {
                const double t = x[gt.gi[g]] * x[gt.gj[g]];
                w[g] = t + t;
            }
        }
        double zs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double rs = 0.0;
        auto row = [&](const double* l) {
            double p = eps;
#pragma unroll
            for (int g = 0; g < G; g++) p = fma(w[g], l[g], p);
            return sleef_log(p);
        };
        const uint32_t n =
            visit([&](int j, const double* l) { zs[j] = row(l) + zs[j]; }, [&](const double* l) { rs = row(l) + rs; });
        return -(radd8(zs) + rs) / (double)n;
    }

    // allele/grad_hess.rs:107-396 + build_grad_hess 83-105
    __device__ void gh(const double* x, double* g, double* h /* MX stride */) const {
        constexpr Genotypes<A> gt;
        double zg[A][8], zh[G][8], rg[A], rh[G];
#pragma unroll
        for (int a = 0; a < A; a++) {
            rg[a] = 0.0;
#pragma unroll
            for (int j = 0; j < 8; j++) zg[a][j] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < G; q++) {
            rh[q] = 0.0;
#pragma unroll
            for (int j = 0; j < 8; j++) zh[q][j] = 0.0;
        }
        auto row = [&](const double* l, int j) {
            double lx[A];
#pragma unroll
            for (int a = 0; a < A; a++) {
                lx[a] = x[0] * l[gl_index(a, 0)];
#pragma unroll
                for (int b = 1; b < A; b++) lx[a] = fma(x[b], l[gl_index(a, b)], lx[a]);
            }
            double d = eps;
#pragma unroll
            for (int a = 0; a < A; a++) d = fma(x[a], lx[a], d);
            const double inv = 1.0 / d;
            double lxd[A];
#pragma unroll
            for (int a = 0; a < A; a++) lxd[a] = lx[a] * inv;
#pragma unroll
            for (int q = 0; q < G; q++) {
                double& z = j < 0 ? rh[q] : zh[q][j];
                z = fma(2.0, lxd[gt.gi[q]] * lxd[gt.gj[q]], z);
                z = fma(-l[q], inv, z);
            }
#pragma unroll
            for (int a = 0; a < A; a++) {
                double& z = j < 0 ? rg[a] : zg[a][j];
                z = fma(lx[a], inv, z);
            }
        };
        const uint32_t n = visit([&](int j, const double* l) { row(l, j); }, [&](const double* l) { row(l, -1); });
        const double nd = (double)n;
#pragma unroll
        for (int a = 0; a < A; a++) g[a] = -2.0 * (radd8(zg[a]) + rg[a]) / nd;
#pragma unroll
        for (int q = 0; q < G; q++) {
            const double v = 2.0 * (radd8(zh[q]) + rh[q]) / nd;
            h[gt.gi[q] * MX + gt.gj[q]] = v;
            h[gt.gj[q] * MX + gt.gi[q]] = v;
        }
    }

    __device__ uint32_t called() const {
        uint32_t n = 0;
        for (uint32_t i = 0; i < n_samples; i++) {
            double l[G];
            n += rows.load(i, l);
        }
        return n;
    }
};

// allele/mod.rs:100-126: SQP from the uniform start (sqp.rs:49-93 with QP cap
// A + 1, set by the caller in tune), then alleles with 0 < p < maf are zeroed
// and the frequencies renormalised.
template <class Rows, int A>
__device__ void fit_site(const AlleleSite<Rows, A>& site, uint32_t s, const Tune& tune, double maf,
                         double* __restrict__ af_out, int* __restrict__ status) {
    if (site.called() == 0) {  // lanevector.rs:73-76 would panic
#pragma unroll
        for (int a = 0; a < A; a++) af_out[(size_t)A * s + a] = 0.0;
        atomicOr(status, 1);
        return;
    }
    double x0[A], x[A];
#pragma unroll
    for (int a = 0; a < A; a++) x0[a] = 1.0 / (double)A;
    sum_to_one<A>(x0, x);
    double f = site.obj(x);
    uint64_t iter = 0;
    for (; iter < tune.sqp_max_iter; iter++) {
        double g[A], h[MX * A];
        site.gh(x, g, h);
        if (converged<A>(x, g, tune.sqp_conv_tol)) break;
        double c[A], y[A];
        for (int i = 0; i < A; i++) c[i] = g[i] - dotn(A, h + i * MX, x);
        solve_qp<A>(h, c, x, tune, y);
        double p[A];
        for (int i = 0; i < A; i++) p[i] = y[i] - x[i];
        const double t = tune.bls_sufficient_decrease * dotn(A, g, p);
        double alpha = 1.0, xn[A], fn;
        uint64_t bi = 0;
        for (;;) {
            for (int i = 0; i < A; i++) xn[i] = x[i] + alpha * p[i];
            sum_to_one<A>(xn, xn);
            fn = site.obj(xn);
            if (fn <= f + alpha * t) break;
            if (++bi == tune.bls_max_iter) break;
            alpha *= tune.bls_step_size_reduce;
        }
        for (int i = 0; i < A; i++) x[i] = xn[i];
        f = fn;
    }
    for (int i = 0; i < A; i++)
        if (x[i] > 0.0 && x[i] < maf) x[i] = 0.0;
    double o[A];
    sum_to_one<A>(x, o);
    for (int a = 0; a < A; a++) af_out[(size_t)A * s + a] = o[a];
}

}  // namespace

// One site per thread: codes (2-bit sample-major rows) or likelihood rows
// (site-major, V x N x G); af_out holds A frequencies per site.
__global__ __launch_bounds__(64) void af_kernel(const uint8_t* __restrict__ packed, size_t row_bytes,
                                                 uint32_t n_samples, uint32_t n_sites, Tune tune, double maf,
                                                 double* __restrict__ af_out, int* __restrict__ status) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sites) return;
    const AlleleSite<CodeRows, 2> site{{packed + (s >> 2), row_bytes, (s & 3u) * 2u}, n_samples, tune.epsilon};
    fit_site(site, s, tune, maf, af_out, status);
}

template <int A>
__global__ __launch_bounds__(64) void af_gl_kernel(const double* __restrict__ gl, uint32_t n_samples,
                                                    uint32_t n_sites, Tune tune, double maf,
                                                    double* __restrict__ af_out, int* __restrict__ status) {
    constexpr int G = A * (A + 1) / 2;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sites) return;
    const AlleleSite<GlRows<G>, A> site{{gl + (size_t)s * n_samples * G}, n_samples, tune.epsilon};
    fit_site(site, s, tune, maf, af_out, status);
}

// Repack kept sites: dst word w of sample i holds kept sites 16w..16w+15.
__global__ __launch_bounds__(256) void gather_kernel(const uint8_t* __restrict__ src, size_t src_row_bytes,
                                                      const uint32_t* __restrict__ kept, uint32_t n_kept,
                                                      uint32_t n_samples, uint32_t* __restrict__ dst,
                                                      uint32_t dst_row_words,
                                                      const uint32_t* __restrict__ dst_rows) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_samples * dst_row_words) return;
    const uint32_t i = (uint32_t)(tid / dst_row_words), w = (uint32_t)(tid % dst_row_words);
    const uint8_t* row = src + (size_t)i * src_row_bytes;
    uint32_t word = 0;
    for (int j = 0; j < 16; j++) {
        const uint32_t t = w * 16 + j;
        uint32_t c = 3;
        if (t < n_kept) {
            const uint32_t s = kept[t];
            c = (row[s >> 2] >> ((s & 3u) * 2u)) & 3u;
        }
        word |= c << (2 * j);
    }
    const size_t drow = dst_rows ? dst_rows[i] : i;
    dst[drow * dst_row_words + w] = word;
}

// ------------------------------------------------------------ the sweep

namespace {

// One site of one pair: L = L0 + c.D as an fma chain over columns 0..8
// (objective.rs:58-63; L0 = eps, or 1.0 for a missing site whose c is 0),
// the log product (the sum of log L over a 16-site group is the log of the
// product of its L; a missing site multiplies by exactly 1.0), r = 1/L
// (grad_hess.rs:122), g += c r, c' = c r, H += c' c'^T (upper 45).
__device__ __forceinline__ void site_update(double* cv, double L0, const double* D, int k, double& P, double& F,
                                            double* G, double* H) {
    double L = L0;
#pragma unroll
    for (int j = 0; j < 9; j++) L = fma(D[j], cv[j], L);
    P = P * L;
    if ((k & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
        F = F + sleef_log_pos(P);
        P = 1.0;
    }
    const double r = recip_rn(L);
#pragma unroll
    for (int j = 0; j < 9; j++) G[j] = fma(cv[j], r, G[j]);
#pragma unroll
    for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
    int id = 0;
#pragma unroll
    for (int a = 0; a < 9; a++)
#pragma unroll
        for (int bb = a; bb < 9; bb++, id++) H[id] = fma(cv[a], cv[bb], H[id]);
}

// Table rows: row = min(4*gx + gy, 11) for the 2-bit codes gx, gy of a site;
// rows 0,1,2,4,5,6,8,9,10 are the nine called genotype pairs, rows 3, 7 and 11
// (a missing call on either side) stay zero.  A row holds the nine c_s values
// and, in slot 9, 1.0 marking a usable site.
constexpr int LUT_ROWS = 12;
constexpr int LUT_W = 10;
// Row stride padded by 16 bytes: rows of one site then start in different
// LDS banks, so lanes reading different genotype-pair rows of the same site
// do not conflict (an unpadded stride of 2560 bytes puts all 12 rows on the
// same banks).
constexpr int LUT_RS = CHUNK * LUT_W + 2;
constexpr int LUT_BUF = LUT_ROWS * LUT_RS;  // doubles per buffer

// Rows of the c_s table for site (p, q): the non-zero block of mixture.rs
// mixture_row2 for one-hot likelihoods of genotype pair (gx, gy), written in
// the same expression order, so each row equals what mixture_row2 returns.
// The block of genotype pair `part` (row-major over (gx, gy) in 0..2) into
// r[9]; returns its table row.
__device__ __forceinline__ int lut_block(double p, double q, int part, double* r) {
    const double p2 = p * p, p3 = p2 * p, q2 = q * q, q3 = q2 * q;
    const double pq = p * q, p2q = p2 * q, pq2 = p * q2, p2q2 = p2 * q2;
    const double p3q = p3 * q, pq3 = p * q3;
    const double twopq = pq + pq, twop2q = p2q + p2q, twop3q = p3q + p3q, twopq2 = pq2 + pq2, twopq3 = pq3 + pq3;
#pragma unroll
    for (int k = 0; k < 9; k++) r[k] = 0.0;
    int row;
    switch (part) {
    case 0:  // (0,0)
        row = 0;
        r[0] = p; r[1] = p2; r[2] = p2; r[3] = p3; r[4] = p2; r[5] = p3; r[6] = p2; r[7] = p3; r[8] = p2 * p2;
        break;
    case 1:  // (0,1)
        row = 1;
        r[2] = pq; r[3] = twop2q; r[7] = p2q; r[8] = twop3q;
        break;
    case 2:  // (0,2)
        row = 2;
        r[1] = pq; r[3] = pq2; r[5] = p2q; r[8] = p2q2;
        break;
    case 3:  // (1,0)
        row = 4;
        r[4] = pq; r[5] = twop2q; r[7] = p2q; r[8] = twop3q;
        break;
    case 4:  // (1,1)
        row = 5;
        r[6] = twopq; r[7] = pq; r[8] = twopq * twopq;
        break;
    case 5:  // (1,2)
        row = 6;
        r[4] = pq; r[5] = twopq2; r[7] = pq2; r[8] = twopq3;
        break;
    case 6:  // (2,0)
        row = 8;
        r[1] = pq; r[3] = p2q; r[5] = pq2; r[8] = p2q2;
        break;
    case 7:  // (2,1)
        row = 9;
        r[2] = pq; r[3] = twopq2; r[7] = pq2; r[8] = twopq3;
        break;
    default:  // (2,2)
        row = 10;
        r[0] = q; r[1] = q2; r[2] = q2; r[3] = q3; r[4] = q2; r[5] = q3; r[6] = q2; r[7] = q3; r[8] = q2 * q2;
        break;
    }
    return row;
}

__device__ __forceinline__ void lut_rows(double p, double q, int part, double* buf, int site, double eps) {
    double r[9];
    const int row = lut_block(p, q, part, r);
    double* dst = buf + row * LUT_RS + site * LUT_W;
#pragma unroll
    for (int k = 0; k < 9; k++) dst[k] = r[k];
    dst[9] = eps;  // start of the L = eps + c.D chain
}

// The mixed-precision form's table: the same rows rounded to FP32, 12 floats
// an entry (c[9], the start of the L chain, 2 of padding: three 16-byte
// reads), row stride padded by 16 bytes as above.
constexpr int LUTF_W = 12;
constexpr int LUTF_RS = CHUNK * LUTF_W + 4;
constexpr int LUTF_BUF = LUT_ROWS * LUTF_RS;  // floats per buffer

__device__ __forceinline__ void lut_rows_f32(double p, double q, int part, float* buf, int site, double eps) {
    double r[9];
    const int row = lut_block(p, q, part, r);
    float* dst = buf + row * LUTF_RS + site * LUTF_W;
#pragma unroll
    for (int k = 0; k < 9; k++) dst[k] = (float)r[k];
    dst[9] = (float)eps;
}

}  // namespace

// One evaluation pass.  blockIdx.y = site segment, blockIdx.x * 256 +
// threadIdx.x = slot (one pair per lane).  Partials (f, g[9], h[45], n) go to
// partial[(seg * NPART + k) * nslots + slot].
// SKIP: a wavefront whose live lanes all pair the same sample x passes over
// the sites where x's call is missing.  Every lane's table row there is a
// missing row (c = 0, L = 1): the site multiplies the log product by 1 and
// adds fma(0, 1, s) = s to the gradient and Hessian sums (which start at +0
// and only gain non-negative terms, so never hold -0), so passing over it
// changes no bit; the log of a 16-site group is still taken at its last site.
template <bool SKIP>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_fp64(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                            const double* __restrict__ af,
                                                            const uint32_t* __restrict__ seg_bounds,
                                                            const uint32_t* __restrict__ sx,
                                                            const uint32_t* __restrict__ sy,
                                                            const double* __restrict__ pt,
                                                            const int* __restrict__ active, uint32_t nslots,
                                                            double eps, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) double lut[2 * LUT_BUF];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    // rows of a missing call on either side (3, 7, 11) stay c = 0 with 1.0
    // as the start of the L chain, so such a site has L = 1 exactly and adds
    // nothing to the log product, the gradient or the Hessian
    for (uint32_t i = tid; i < 2 * LUT_BUF; i += SWEEP_TPB) {
        const uint32_t off = i % LUT_BUF, row = off / LUT_RS, w = off % LUT_RS;
        lut[i] = (row & 3u) == 3u && w < CHUNK * LUT_W && w % LUT_W == 9 ? 1.0 : 0.0;
    }
    __syncthreads();

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    // SKIP: the first live lane, and whether every live lane has its x
    bool x_shared = false;
    int x_lane = 0;
    if (SKIP) {
        const unsigned long long lm = __ballot(live);
        if (lm) {
            x_lane = __ffsll((long long)lm) - 1;
            const uint32_t xf = (uint32_t)__builtin_amdgcn_readlane((int)x, x_lane);
            x_shared = __all(!live || x == xf);
        }
    }

    double F = 0.0, N = 0.0, P = 1.0, G[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;

    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    // (site, row) of the table rows this thread builds: the 32 sites of a chunk
    // across the lanes and the rows across the wavefronts, so that a
    // wavefront's lanes take two of lut_block's cases (three for wavefront 0,
    // which adds row (2,2)) instead of all nine
    const int lut_site = tid & (CHUNK - 1), lut_part = tid / CHUNK;
    auto build = [&](double p, double q, int b) {
        double* buf = lut + b * LUT_BUF;
        lut_rows(p, q, lut_part, buf, lut_site, eps);
        if (lut_part == 0) lut_rows(p, q, 8, buf, lut_site, eps);
    };
    // the genotype words of a chunk and the frequencies of the LUT built for
    // the next one are loaded a chunk ahead of their use
    uint2 wx = make_uint2(0u, 0u), wy = make_uint2(0u, 0u);
    double pn = 0.5, qn = 0.5;
    if (s0 < s1) {
        build(af[2 * (s0 + lut_site)], af[2 * (s0 + lut_site) + 1], 0);
        wx = *reinterpret_cast<const uint2*>(rowx + s0 / 16);
        wy = *reinterpret_cast<const uint2*>(rowy + s0 / 16);
    }
    if (s0 + CHUNK < s1) {
        pn = af[2 * (s0 + CHUNK + lut_site)];
        qn = af[2 * (s0 + CHUNK + lut_site) + 1];
    }
    __syncthreads();
    int b = 0;
    for (uint32_t cs = s0; cs < s1; cs += CHUNK, b ^= 1) {
        uint2 nwx = wx, nwy = wy;
        double pnn = 0.5, qnn = 0.5;
        if (cs + CHUNK < s1) {
            nwx = *reinterpret_cast<const uint2*>(rowx + (cs + CHUNK) / 16);
            nwy = *reinterpret_cast<const uint2*>(rowy + (cs + CHUNK) / 16);
            build(pn, qn, b ^ 1);
        }
        if (cs + 2 * CHUNK < s1) {
            pnn = af[2 * (cs + 2 * CHUNK + lut_site)];
            qnn = af[2 * (cs + 2 * CHUNK + lut_site) + 1];
        }
        const double* L0 = lut + b * LUT_BUF;
        // usable sites of the chunk: neither call missing (code 3), counted
        // exactly as integers (lanevector.rs:100-104 NaN mask)
        {
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
        }
        // SKIP: the chunk's sites where the shared x is missing (bit 2k of
        // the word of sites 16h..16h+15), uniform over the wavefront
        uint32_t skip_lo = 0, skip_hi = 0;
        if (SKIP && x_shared) {
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)wx.x, x_lane);
            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)wx.y, x_lane);
            skip_lo = a & (a >> 1) & 0x55555555u;
            skip_hi = c & (c >> 1) & 0x55555555u;
        }
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            if (SKIP && (((k < 16 ? skip_lo : skip_hi) >> (2 * (k & 15))) & 1u)) {
                if ((k & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
                    F = F + sleef_log_pos(P);
                    P = 1.0;
                }
                continue;
            }
            const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
            const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
            const uint32_t row = min(((ax & 3u) << 2) | (ay & 3u), 11u);
            const double2* c2 = reinterpret_cast<const double2*>(L0 + row * LUT_RS + k * LUT_W);
            double cv[10];
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const double2 v = c2[j];
                cv[2 * j] = v.x;
                cv[2 * j + 1] = v.y;
            }
            // column 9 holds eps, or 1.0 on a missing row
            site_update(cv, cv[9], D, k, P, F, G, H);
        }
        wx = nwx;
        wy = nwy;
        pn = pnn;
        qn = qnn;
        __syncthreads();
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = G[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
        out[(size_t)55 * nslots] = N;
    }
}

// c_s of one site for the 2-bit codes (ax, ay) (both called), built in
// registers from the site's frequencies: the value lut_block writes into
// table row 4*ax + ay, bit for bit.  Swapping p and q maps genotype pair
// (gx, gy) to (2 - gx, 2 - gy), and every product of lut_block to the same
// two operands in the other order (p2q of (q, p) is q2*p = pq2 of (p, q), and
// so on), so the five pairs with 3*ax + ay <= 4 cover the nine; each
// component is then a select among values computed once.
__device__ __forceinline__ void cs_from_codes(double p, double q, uint32_t ax, uint32_t ay, double* r) {
    int part = (int)(3 * ax + ay);
    const bool sw = part > 4;
    const double a = sw ? q : p, b = sw ? p : q;
    part = sw ? 8 - part : part;
    const double a2 = a * a, a3 = a2 * a, b2 = b * b, ab = a * b;
    const double a2b = a2 * b, ab2 = a * b2, a2b2 = a2 * b2, a3b = a3 * b;
    const double tab = ab + ab, ta2b = a2b + a2b, ta3b = a3b + a3b;
    const double a4 = a2 * a2, tab2 = tab * tab;
    const bool p0 = part == 0, p1 = part == 1, p2 = part == 2, p3 = part == 3, p4 = part == 4;
    r[0] = p0 ? a : 0.0;
    r[1] = p0 ? a2 : p2 ? ab : 0.0;
    r[2] = p0 ? a2 : p1 ? ab : 0.0;
    r[3] = p0 ? a3 : p1 ? ta2b : p2 ? ab2 : 0.0;
    r[4] = p0 ? a2 : p3 ? ab : 0.0;
    r[5] = p0 ? a3 : p2 ? a2b : p3 ? ta2b : 0.0;
    r[6] = p0 ? a2 : p4 ? tab : 0.0;
    r[7] = p0 ? a3 : (p1 || p3) ? a2b : p4 ? ab : 0.0;
    r[8] = p0 ? a4 : (p1 || p3) ? ta3b : p2 ? a2b2 : tab2;
}

// One usable site of the wavefront form in a lane: L = eps + c.D, the log
// product's factor, r = 1/L, g += c r, c' = c r, H += c' c'^T (site_update's
// operations).
__device__ __forceinline__ void lanes_site(double* cv, double eps, const double* D, double& P, double* G,
                                           double* H) {
    double L = eps;
#pragma unroll
    for (int j = 0; j < 9; j++) L = fma(D[j], cv[j], L);
    P = P * L;
    const double r = recip_rn(L);
#pragma unroll
    for (int j = 0; j < 9; j++) G[j] = fma(cv[j], r, G[j]);
#pragma unroll
    for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
    int id = 0;
#pragma unroll
    for (int a = 0; a < 9; a++)
#pragma unroll
        for (int bb = a; bb < 9; bb++, id++) H[id] = fma(cv[a], cv[bb], H[id]);
}

// The 64 lanes' sums added in a butterfly (xor 32, 16, ..., 1: lane i adds
// lane i ^ o's value; both partners hold the same sum since IEEE addition
// commutes); lane 0 writes the segment's partials (out = the slot's entry of
// segment 0 of them, stride nslots).
__device__ __forceinline__ void lanes_finish(double F, double N, double* G, double* H, uint32_t lane, double* out,
                                             uint32_t nslots) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        F = F + __shfl_xor(F, o, 64);
        N = N + __shfl_xor(N, o, 64);
#pragma unroll
        for (int k = 0; k < 9; k++) G[k] = G[k] + __shfl_xor(G[k], o, 64);
#pragma unroll
        for (int k = 0; k < 45; k++) H[k] = H[k] + __shfl_xor(H[k], o, 64);
    }
    if (lane == 0) {
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = G[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
        out[(size_t)55 * nslots] = N;
    }
}

// The sweep with a pair's sites across the lanes of a wavefront (option
// "pair_lanes" 64), for fits with fewer pairs than the device has lanes: one
// pair per wavefront, four a workgroup (wavefront w of block b takes slot
// 4b + w), blockIdx.y = site segment.  Lane l takes the sites s0 + l + 64 t
// of its segment, t = 0, 1, ..., builds c_s in registers from the 2-bit codes
// and the site's frequencies (cs_from_codes), and keeps its own sums in site
// order: the log product over its usable sites with a log after every 16 of
// its steps (and after its last step), the gradient and the Hessian as in
// site_update.  A site with a missing call is passed over (c = 0, L = 1 adds
// nothing: the argument of sweep_fp64<true>).  At the segment's end the 64
// lanes' sums are added in a butterfly (xor 32, 16, ..., 1: lane i adds lane
// i ^ o's value, and both partners hold the same sum since IEEE addition
// commutes), and lane 0 writes the partials in sweep_fp64's layout, so the
// fold and the step are unchanged.  The summation order differs from
// sweep_fp64's, so records differ from the one-pair-per-lane form's in the
// last bits; the oracle restates this order.
__global__ __launch_bounds__(SWEEP_TPB) void sweep_lanes_fp64(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                               const double* __restrict__ af,
                                                               const uint32_t* __restrict__ seg_bounds,
                                                               const uint32_t* __restrict__ sx,
                                                               const uint32_t* __restrict__ sy,
                                                               const double* __restrict__ pt,
                                                               const int* __restrict__ active, uint32_t nslots,
                                                               double eps, double* __restrict__ partial) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t slot = blockIdx.x * (SWEEP_TPB / 64) +
                          (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (slot >= nslots || active[slot] == 0) return;  // uniform over the wavefront
    const uint32_t seg = blockIdx.y;
    const uint32_t x = sx[slot], y = sy[slot];
    double D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];

    double F = 0.0, N = 0.0, P = 1.0, G[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;
    const uint32_t steps = (s1 - s0 + 63) / 64;
    for (uint32_t t = 0; t < steps; t++) {
        const uint32_t s = s0 + 64 * t + lane;
        if (s < s1) {
            const uint32_t ax = (rowx[s >> 4] >> (2 * (s & 15u))) & 3u;
            const uint32_t ay = (rowy[s >> 4] >> (2 * (s & 15u))) & 3u;
            if (ax != 3u && ay != 3u) {
                const double2 f = *reinterpret_cast<const double2*>(af + 2 * (size_t)s);
                double cv[9];
                cs_from_codes(f.x, f.y, ax, ay, cv);
                lanes_site(cv, eps, D, P, G, H);
                N += 1.0;
            }
        }
        if ((t & (LOG_GROUP - 1)) == LOG_GROUP - 1 || t + 1 == steps) {
            F = F + sleef_log_pos(P);
            P = 1.0;
        }
    }
    lanes_finish(F, N, G, H, lane, partial + (size_t)seg * NPART * nslots + slot, nslots);
}

// The mixed-precision form (precision KESTREL_MIXED): the pass of sweep_fp64
// with the c_s table in FP32 and the Hessian in FP32 -- 1/L by v_rcp_f32, the
// scaled c' = c/L and the 45 Hessian terms, summed in FP32 over the segment
// (without FP64 Hessian totals the sweep needs 168 VGPRs: three wavefronts
// per SIMD; the Hessian only steers the Newton step, and fits take the FP64
// form's iterations).  L = eps + c.D and the gradient terms c * (1/L) are
// FP64 (from the FP32 c and 1/L, whose products are exact in FP64): the line
// search compares objectives, and the convergence test |g_i - lambda| <= 1e-8
// reads the gradient, at resolutions FP32 sums do not reach (with L or the
// gradient in FP32 the fits took 1.35x the iterations and 3.4x the
// evaluations).  The log product, the objective and everything the step does
// stay FP64.  Results are not the FP64 form's bits.  With u = 2^-24 and a
// segment of n sites, c' = c (1/L) carries a relative error of at most 4u
// (c rounded, v_rcp_f32 within 1 ulp, the product rounded), so each Hessian
// term is within (8 + n) u sqrt(H_ii H_jj) of the exact one: ||dH|| <= 9 (8 +
// n) u ||H|| (2.2e-3 at n = 4096), which moves a log10 condition below 100 by
// less than 0.1; larger conditions are estimated again from an FP64 pass
// (RECOND_MIN in the host code).  SKIP: sites where the wavefront's shared sample x is
// missing are passed over as in sweep_fp64<true> (c = 0 and L = 1 there, so
// no FP32 or FP64 sum changes: the same bits as without).
template <bool SKIP>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_mixed(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                             const double* __restrict__ af,
                                                             const uint32_t* __restrict__ seg_bounds,
                                                             const uint32_t* __restrict__ sx,
                                                             const uint32_t* __restrict__ sy,
                                                             const double* __restrict__ pt,
                                                             const int* __restrict__ active, uint32_t nslots,
                                                             double eps, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) float lut[2 * LUTF_BUF];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    // missing rows (3, 7, 11): c = 0 and 1.0 as the start of the L chain
    for (uint32_t i = tid; i < 2 * LUTF_BUF; i += SWEEP_TPB) {
        const uint32_t off = i % LUTF_BUF, row = off / LUTF_RS, w = off % LUTF_RS;
        lut[i] = (row & 3u) == 3u && w < CHUNK * LUTF_W && w % LUTF_W == 9 ? 1.0f : 0.0f;
    }
    __syncthreads();

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    bool x_shared = false;
    int x_lane = 0;
    if (SKIP) {
        const unsigned long long lm = __ballot(live);
        if (lm) {
            x_lane = __ffsll((long long)lm) - 1;
            const uint32_t xf = (uint32_t)__builtin_amdgcn_readlane((int)x, x_lane);
            x_shared = __all(!live || x == xf);
        }
    }

    double F = 0.0, N = 0.0, P = 1.0, G[9];
    float h[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) h[k] = 0.0f;

    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    // (site, row) of the table rows this thread builds: the 32 sites of a chunk
    // across the lanes and the rows across the wavefronts, so that a
    // wavefront's lanes take two of lut_block's cases (three for wavefront 0,
    // which adds row (2,2)) instead of all nine
    const int lut_site = tid & (CHUNK - 1), lut_part = tid / CHUNK;
    auto build = [&](double p, double q, int b) {
        float* buf = lut + b * LUTF_BUF;
        lut_rows_f32(p, q, lut_part, buf, lut_site, eps);
        if (lut_part == 0) lut_rows_f32(p, q, 8, buf, lut_site, eps);
    };
    uint2 wx = make_uint2(0u, 0u), wy = make_uint2(0u, 0u);
    double pn = 0.5, qn = 0.5;
    if (s0 < s1) {
        build(af[2 * (s0 + lut_site)], af[2 * (s0 + lut_site) + 1], 0);
        wx = *reinterpret_cast<const uint2*>(rowx + s0 / 16);
        wy = *reinterpret_cast<const uint2*>(rowy + s0 / 16);
    }
    if (s0 + CHUNK < s1) {
        pn = af[2 * (s0 + CHUNK + lut_site)];
        qn = af[2 * (s0 + CHUNK + lut_site) + 1];
    }
    __syncthreads();
    int b = 0;
    for (uint32_t cs = s0; cs < s1; cs += CHUNK, b ^= 1) {
        uint2 nwx = wx, nwy = wy;
        double pnn = 0.5, qnn = 0.5;
        if (cs + CHUNK < s1) {
            nwx = *reinterpret_cast<const uint2*>(rowx + (cs + CHUNK) / 16);
            nwy = *reinterpret_cast<const uint2*>(rowy + (cs + CHUNK) / 16);
            build(pn, qn, b ^ 1);
        }
        if (cs + 2 * CHUNK < s1) {
            pnn = af[2 * (cs + 2 * CHUNK + lut_site)];
            qnn = af[2 * (cs + 2 * CHUNK + lut_site) + 1];
        }
        const float* L0 = lut + b * LUTF_BUF;
        {
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
        }
        uint32_t skip_lo = 0, skip_hi = 0;
        if (SKIP && x_shared) {
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)wx.x, x_lane);
            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)wx.y, x_lane);
            skip_lo = a & (a >> 1) & 0x55555555u;
            skip_hi = c & (c >> 1) & 0x55555555u;
        }
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            if (SKIP && (((k < 16 ? skip_lo : skip_hi) >> (2 * (k & 15))) & 1u)) {
                if ((k & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
                    F = F + sleef_log_pos(P);
                    P = 1.0;
                }
                continue;
            }
            const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
            const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
            const uint32_t row = min(((ax & 3u) << 2) | (ay & 3u), 11u);
            const float4* c4 = reinterpret_cast<const float4*>(L0 + row * LUTF_RS + k * LUTF_W);
            const float4 u = c4[0], v = c4[1], w = c4[2];
            float cv[9] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w, w.x};
            double cd[9];
#pragma unroll
            for (int j = 0; j < 9; j++) cd[j] = (double)cv[j];
            double L = w.y == 1.0f ? 1.0 : eps;  // 1.0 on a missing row
#pragma unroll
            for (int j = 0; j < 9; j++) L = fma(D[j], cd[j], L);
            P = P * L;
            if ((k & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
                F = F + sleef_log_pos(P);
                P = 1.0;
            }
            const float r = __builtin_amdgcn_rcpf((float)L);
            const double rd = (double)r;
#pragma unroll
            for (int j = 0; j < 9; j++) G[j] = fma(cd[j], rd, G[j]);
#pragma unroll
            for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
            int id = 0;
#pragma unroll
            for (int a2 = 0; a2 < 9; a2++)
#pragma unroll
                for (int bb = a2; bb < 9; bb++, id++) h[id] = fmaf(cv[a2], cv[bb], h[id]);
        }
        wx = nwx;
        wy = nwy;
        pn = pnn;
        qn = qnn;
        __syncthreads();
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = G[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = (double)h[k];
        out[(size_t)55 * nslots] = N;
    }
}

// Soft likelihoods: the same pass with c_s built per pair and site from the two
// likelihood triples in the operation order of mixture_row2 (mixture.rs:96-236):
// 9 products lx[i] ly[j], then 45 fma into c (each block (i, j) in the
// reference's order) with the site's 18 frequency products, which the
// workgroup computes once per site into LDS (K_*).  gl holds sample-major rows
// of v_pad triples (0 for a missing call); the 2-bit mask rows (code 3 =
// missing) select the missing sites and count n_xy.
enum : int {
    K_P, K_P2, K_P3, K_P4, K_Q, K_Q2, K_Q3, K_Q4, K_PQ, K_2P2Q, K_P2Q, K_2P3Q, K_PQ2, K_P2Q2, K_2PQ, K_2PQSQ,
    K_2PQ2, K_2PQ3, K_N
};

__device__ __forceinline__ void site_coefficients(double p, double q, double* K) {
    const double p2 = p * p, p3 = p2 * p, q2 = q * q, q3 = q2 * q;
    const double pq = p * q, p2q = p2 * q, pq2 = p * q2, p2q2 = p2 * q2;
    const double p3q = p3 * q, pq3 = p * q3;
    const double twopq = pq + pq, twop2q = p2q + p2q, twop3q = p3q + p3q, twopq2 = pq2 + pq2, twopq3 = pq3 + pq3;
    K[K_P] = p;
    K[K_P2] = p2;
    K[K_P3] = p3;
    K[K_P4] = p2 * p2;
    K[K_Q] = q;
    K[K_Q2] = q2;
    K[K_Q3] = q3;
    K[K_Q4] = q2 * q2;
    K[K_PQ] = pq;
    K[K_2P2Q] = twop2q;
    K[K_P2Q] = p2q;
    K[K_2P3Q] = twop3q;
    K[K_PQ2] = pq2;
    K[K_P2Q2] = p2q2;
    K[K_2PQ] = twopq;
    K[K_2PQSQ] = twopq * twopq;
    K[K_2PQ2] = twopq2;
    K[K_2PQ3] = twopq3;
}

// The mixed-precision form's coefficients: the FP64 ones rounded to FP32,
// KF_W floats a site (K_N and 2 of padding: 16-byte reads).
constexpr int KF_W = 20;
constexpr int KF_BUF = CHUNK * KF_W;  // floats per coefficient buffer

__device__ __forceinline__ void site_coefficients_f32(double p, double q, float* K) {
    double Kd[K_N];
    site_coefficients(p, q, Kd);
#pragma unroll
    for (int k = 0; k < K_N; k++) K[k] = (float)Kd[k];
}

// mixture.rs:138-233 block by block (T = float in the mixed-precision form)
template <typename T>
__device__ __forceinline__ void mixture_soft(const T* lx, const T* ly, const T* K, T* a) {
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = T(0);
    T g = lx[0] * ly[0];  // (0,0) (0,0)
    a[0] = fma(g, K[K_P], a[0]);
    a[1] = fma(g, K[K_P2], a[1]);
    a[2] = fma(g, K[K_P2], a[2]);
    a[3] = fma(g, K[K_P3], a[3]);
    a[4] = fma(g, K[K_P2], a[4]);
    a[5] = fma(g, K[K_P3], a[5]);
    a[6] = fma(g, K[K_P2], a[6]);
    a[7] = fma(g, K[K_P3], a[7]);
    a[8] = fma(g, K[K_P4], a[8]);
    g = lx[0] * ly[1];  // (0,0) (0,1)
    a[2] = fma(g, K[K_PQ], a[2]);
    a[3] = fma(g, K[K_2P2Q], a[3]);
    a[7] = fma(g, K[K_P2Q], a[7]);
    a[8] = fma(g, K[K_2P3Q], a[8]);
    g = lx[0] * ly[2];  // (0,0) (1,1)
    a[1] = fma(g, K[K_PQ], a[1]);
    a[3] = fma(g, K[K_PQ2], a[3]);
    a[5] = fma(g, K[K_P2Q], a[5]);
    a[8] = fma(g, K[K_P2Q2], a[8]);
    g = lx[1] * ly[0];  // (0,1) (0,0)
    a[4] = fma(g, K[K_PQ], a[4]);
    a[5] = fma(g, K[K_2P2Q], a[5]);
    a[7] = fma(g, K[K_P2Q], a[7]);
    a[8] = fma(g, K[K_2P3Q], a[8]);
    g = lx[1] * ly[1];  // (0,1) (0,1)
    a[6] = fma(g, K[K_2PQ], a[6]);
    a[7] = fma(g, K[K_PQ], a[7]);
    a[8] = fma(g, K[K_2PQSQ], a[8]);
    g = lx[1] * ly[2];  // (0,1) (1,1)
    a[4] = fma(g, K[K_PQ], a[4]);
    a[5] = fma(g, K[K_2PQ2], a[5]);
    a[7] = fma(g, K[K_PQ2], a[7]);
    a[8] = fma(g, K[K_2PQ3], a[8]);
    g = lx[2] * ly[0];  // (1,1) (0,0)
    a[1] = fma(g, K[K_PQ], a[1]);
    a[3] = fma(g, K[K_P2Q], a[3]);
    a[5] = fma(g, K[K_PQ2], a[5]);
    a[8] = fma(g, K[K_P2Q2], a[8]);
    g = lx[2] * ly[1];  // (1,1) (0,1)
    a[2] = fma(g, K[K_PQ], a[2]);
    a[3] = fma(g, K[K_2PQ2], a[3]);
    a[7] = fma(g, K[K_PQ2], a[7]);
    a[8] = fma(g, K[K_2PQ3], a[8]);
    g = lx[2] * ly[2];  // (1,1) (1,1)
    a[0] = fma(g, K[K_Q], a[0]);
    a[1] = fma(g, K[K_Q2], a[1]);
    a[2] = fma(g, K[K_Q2], a[2]);
    a[3] = fma(g, K[K_Q3], a[3]);
    a[4] = fma(g, K[K_Q2], a[4]);
    a[5] = fma(g, K[K_Q3], a[5]);
    a[6] = fma(g, K[K_Q2], a[6]);
    a[7] = fma(g, K[K_Q3], a[7]);
    a[8] = fma(g, K[K_Q4], a[8]);
}

constexpr int KT_BUF = CHUNK * K_N;  // doubles per coefficient buffer

// STRIDE 3: rows of three likelihoods (24 bytes per sample and site); STRIDE
// 2: 16-byte rows of the two entries beside an exact 1.0 whose position is the
// site's 2-bit code (gather_gl_kernel), rebuilt here into the same triple.
template <int STRIDE>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_gl(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                          const double* __restrict__ gl, uint32_t v_pad,
                                                          const double* __restrict__ af,
                                                          const uint32_t* __restrict__ seg_bounds,
                                                          const uint32_t* __restrict__ sx,
                                                          const uint32_t* __restrict__ sy,
                                                          const double* __restrict__ pt,
                                                          const int* __restrict__ active, uint32_t nslots,
                                                          double eps, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) double kt[2 * KT_BUF];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const double* glx = gl + (size_t)x * v_pad * STRIDE;
    const double* gly = gl + (size_t)y * v_pad * STRIDE;

    double F = 0.0, N = 0.0, P = 1.0, G[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;

    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    auto build = [&](uint32_t cs, int b) {
        if (tid < CHUNK) site_coefficients(af[2 * (cs + tid)], af[2 * (cs + tid) + 1], kt + b * KT_BUF + tid * K_N);
    };
    if (s0 < s1) build(s0, 0);
    __syncthreads();
    int b = 0;
    for (uint32_t cs = s0; cs < s1; cs += CHUNK, b ^= 1) {
        if (cs + CHUNK < s1) build(cs + CHUNK, b ^ 1);
        const uint2 wx = *reinterpret_cast<const uint2*>(rowx + cs / 16);
        const uint2 wy = *reinterpret_cast<const uint2*>(rowy + cs / 16);
        {
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
        }
        const double* K0 = kt + b * KT_BUF;
#pragma unroll
        for (int k = 0; k < CHUNK; k += 2) {
            double lx[2][3], ly[2][3];
            if (STRIDE == 3) {
                // two sites of likelihood triples = 48 bytes, 16-byte aligned
                const double2* px = reinterpret_cast<const double2*>(glx + (size_t)(cs + k) * 3);
                const double2* py = reinterpret_cast<const double2*>(gly + (size_t)(cs + k) * 3);
                const double2 x0 = px[0], x1 = px[1], x2 = px[2];
                const double2 y0 = py[0], y1 = py[1], y2 = py[2];
                lx[0][0] = x0.x; lx[0][1] = x0.y; lx[0][2] = x1.x;
                lx[1][0] = x1.y; lx[1][1] = x2.x; lx[1][2] = x2.y;
                ly[0][0] = y0.x; ly[0][1] = y0.y; ly[0][2] = y1.x;
                ly[1][0] = y1.y; ly[1][1] = y2.x; ly[1][2] = y2.y;
            } else {
                // two sites of 16-byte rows: put the exact 1.0 back at its position
                const double2* px = reinterpret_cast<const double2*>(glx + (size_t)(cs + k) * 2);
                const double2* py = reinterpret_cast<const double2*>(gly + (size_t)(cs + k) * 2);
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int kk = k + h;
                    const uint32_t cx = ((kk < 16 ? wx.x : wx.y) >> (2 * (kk & 15))) & 3u;
                    const uint32_t cy = ((kk < 16 ? wy.x : wy.y) >> (2 * (kk & 15))) & 3u;
                    const double2 vx = px[h], vy = py[h];
                    lx[h][0] = cx == 0 ? 1.0 : vx.x;
                    lx[h][1] = cx == 0 ? vx.x : cx == 1 ? 1.0 : vx.y;
                    lx[h][2] = cx == 2 ? 1.0 : vx.y;
                    ly[h][0] = cy == 0 ? 1.0 : vy.x;
                    ly[h][1] = cy == 0 ? vy.x : cy == 1 ? 1.0 : vy.y;
                    ly[h][2] = cy == 2 ? 1.0 : vy.y;
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int kk = k + h;
                const uint32_t ax = (kk < 16 ? wx.x : wx.y) >> (2 * (kk & 15));
                const uint32_t ay = (kk < 16 ? wy.x : wy.y) >> (2 * (kk & 15));
                const bool missing = (ax & 3u) == 3u || (ay & 3u) == 3u;
                double K[K_N];
                const double2* k2 = reinterpret_cast<const double2*>(K0 + kk * K_N);
#pragma unroll
                for (int j = 0; j < K_N / 2; j++) {
                    const double2 v = k2[j];
                    K[2 * j] = v.x;
                    K[2 * j + 1] = v.y;
                }
                double cv[9];
                mixture_soft(lx[h], ly[h], K, cv);
                site_update(cv, missing ? 1.0 : eps, D, kk, P, F, G, H);
            }
        }
        __syncthreads();
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = G[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
        out[(size_t)55 * nslots] = N;
    }
}

// The wavefront form of sweep_gl (option "pair_lanes" 64): as
// sweep_lanes_fp64 with likelihood rows -- lane l takes the sites s0 + l +
// 64 t of the segment, builds the site's coefficients from its (p, q) and c_s
// by mixture_soft (the bits of sweep_gl's c_s) from the pair's two rows
// (STRIDE 2: the exact 1.0 put back at the position of the 2-bit code), and
// passes over sites where either call is missing (c = 0 there, L = 1).
template <int STRIDE>
__global__ __launch_bounds__(SWEEP_TPB) void sweep_lanes_gl(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                             const double* __restrict__ gl, uint32_t v_pad,
                                                             const double* __restrict__ af,
                                                             const uint32_t* __restrict__ seg_bounds,
                                                             const uint32_t* __restrict__ sx,
                                                             const uint32_t* __restrict__ sy,
                                                             const double* __restrict__ pt,
                                                             const int* __restrict__ active, uint32_t nslots,
                                                             double eps, double* __restrict__ partial) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t slot = blockIdx.x * (SWEEP_TPB / 64) +
                          (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (slot >= nslots || active[slot] == 0) return;  // uniform over the wavefront
    const uint32_t seg = blockIdx.y;
    const uint32_t x = sx[slot], y = sy[slot];
    double D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const double* glx = gl + (size_t)x * v_pad * STRIDE;
    const double* gly = gl + (size_t)y * v_pad * STRIDE;
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];

    auto load_row = [&](const double* base, uint32_t t, uint32_t code, double* l) {
        if (STRIDE == 3) {
            const double* r = base + (size_t)t * 3;
            l[0] = r[0];
            l[1] = r[1];
            l[2] = r[2];
        } else {
            const double2 v = *reinterpret_cast<const double2*>(base + (size_t)t * 2);
            l[0] = code == 0 ? 1.0 : v.x;
            l[1] = code == 0 ? v.x : code == 1 ? 1.0 : v.y;
            l[2] = code == 2 ? 1.0 : v.y;
        }
    };
    double F = 0.0, N = 0.0, P = 1.0, G[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;
    const uint32_t steps = (s1 - s0 + 63) / 64;
    for (uint32_t t = 0; t < steps; t++) {
        const uint32_t s = s0 + 64 * t + lane;
        if (s < s1) {
            const uint32_t ax = (rowx[s >> 4] >> (2 * (s & 15u))) & 3u;
            const uint32_t ay = (rowy[s >> 4] >> (2 * (s & 15u))) & 3u;
            if (ax != 3u && ay != 3u) {
                const double2 f = *reinterpret_cast<const double2*>(af + 2 * (size_t)s);
                double K[K_N], lx[3], ly[3], cv[9];
                site_coefficients(f.x, f.y, K);
                load_row(glx, s, ax, lx);
                load_row(gly, s, ay, ly);
                mixture_soft(lx, ly, K, cv);
                lanes_site(cv, eps, D, P, G, H);
                N += 1.0;
            }
        }
        if ((t & (LOG_GROUP - 1)) == LOG_GROUP - 1 || t + 1 == steps) {
            F = F + sleef_log_pos(P);
            P = 1.0;
        }
    }
    lanes_finish(F, N, G, H, lane, partial + (size_t)seg * NPART * nslots + slot, nslots);
}

// The mixed-precision form of sweep_gl (precision KESTREL_MIXED): the site
// coefficients, the likelihood products and c_s (mixture_row2's order), 1/L,
// the scaled c' and the 45 Hessian terms in FP32 (the Hessian summed in FP32
// over the segment); L = eps + c.D, the gradient terms (exact FP64 products
// of the FP32 c and 1/L), the log product and the objective in FP64, as in
// sweep_mixed.
template <int STRIDE>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_gl_mixed(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                                const double* __restrict__ gl, uint32_t v_pad,
                                                                const double* __restrict__ af,
                                                                const uint32_t* __restrict__ seg_bounds,
                                                                const uint32_t* __restrict__ sx,
                                                                const uint32_t* __restrict__ sy,
                                                                const double* __restrict__ pt,
                                                                const int* __restrict__ active, uint32_t nslots,
                                                                double eps, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) float kt[2 * KF_BUF];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const double* glx = gl + (size_t)x * v_pad * STRIDE;
    const double* gly = gl + (size_t)y * v_pad * STRIDE;

    // the Hessian is summed in FP32 over the whole segment (its FP64 totals
    // beside FP32 chunk sums would not fit the registers with the rows)
    double F = 0.0, N = 0.0, P = 1.0, G[9];
    float h[45];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) h[k] = 0.0f;

    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    auto build = [&](uint32_t cs, int b) {
        if (tid < CHUNK)
            site_coefficients_f32(af[2 * (cs + tid)], af[2 * (cs + tid) + 1], kt + b * KF_BUF + tid * KF_W);
    };
    if (s0 < s1) build(s0, 0);
    __syncthreads();
    int b = 0;
    for (uint32_t cs = s0; cs < s1; cs += CHUNK, b ^= 1) {
        if (cs + CHUNK < s1) build(cs + CHUNK, b ^ 1);
        const uint2 wx = *reinterpret_cast<const uint2*>(rowx + cs / 16);
        const uint2 wy = *reinterpret_cast<const uint2*>(rowy + cs / 16);
        {
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
        }
        const float* K0 = kt + b * KF_BUF;
#pragma unroll
        for (int k = 0; k < CHUNK; k += 2) {
            double lx[2][3], ly[2][3];
            if (STRIDE == 3) {
                const double2* px = reinterpret_cast<const double2*>(glx + (size_t)(cs + k) * 3);
                const double2* py = reinterpret_cast<const double2*>(gly + (size_t)(cs + k) * 3);
                const double2 x0 = px[0], x1 = px[1], x2 = px[2];
                const double2 y0 = py[0], y1 = py[1], y2 = py[2];
                lx[0][0] = x0.x; lx[0][1] = x0.y; lx[0][2] = x1.x;
                lx[1][0] = x1.y; lx[1][1] = x2.x; lx[1][2] = x2.y;
                ly[0][0] = y0.x; ly[0][1] = y0.y; ly[0][2] = y1.x;
                ly[1][0] = y1.y; ly[1][1] = y2.x; ly[1][2] = y2.y;
            } else {
                const double2* px = reinterpret_cast<const double2*>(glx + (size_t)(cs + k) * 2);
                const double2* py = reinterpret_cast<const double2*>(gly + (size_t)(cs + k) * 2);
#pragma unroll
                for (int hh = 0; hh < 2; hh++) {
                    const int kk = k + hh;
                    const uint32_t cx = ((kk < 16 ? wx.x : wx.y) >> (2 * (kk & 15))) & 3u;
                    const uint32_t cy = ((kk < 16 ? wy.x : wy.y) >> (2 * (kk & 15))) & 3u;
                    const double2 vx = px[hh], vy = py[hh];
                    lx[hh][0] = cx == 0 ? 1.0 : vx.x;
                    lx[hh][1] = cx == 0 ? vx.x : cx == 1 ? 1.0 : vx.y;
                    lx[hh][2] = cx == 2 ? 1.0 : vx.y;
                    ly[hh][0] = cy == 0 ? 1.0 : vy.x;
                    ly[hh][1] = cy == 0 ? vy.x : cy == 1 ? 1.0 : vy.y;
                    ly[hh][2] = cy == 2 ? 1.0 : vy.y;
                }
            }
#pragma unroll
            for (int hh = 0; hh < 2; hh++) {
                const int kk = k + hh;
                const uint32_t ax = (kk < 16 ? wx.x : wx.y) >> (2 * (kk & 15));
                const uint32_t ay = (kk < 16 ? wy.x : wy.y) >> (2 * (kk & 15));
                const bool missing = (ax & 3u) == 3u || (ay & 3u) == 3u;
                float K[KF_W];
                const float4* k4 = reinterpret_cast<const float4*>(K0 + kk * KF_W);
#pragma unroll
                for (int j = 0; j < KF_W / 4; j++) {
                    const float4 v = k4[j];
                    K[4 * j] = v.x;
                    K[4 * j + 1] = v.y;
                    K[4 * j + 2] = v.z;
                    K[4 * j + 3] = v.w;
                }
                float lxf[3], lyf[3], cv[9];
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    lxf[j] = (float)lx[hh][j];
                    lyf[j] = (float)ly[hh][j];
                }
                mixture_soft<float>(lxf, lyf, K, cv);
                double cd[9];
#pragma unroll
                for (int j = 0; j < 9; j++) cd[j] = (double)cv[j];
                double L = missing ? 1.0 : eps;
#pragma unroll
                for (int j = 0; j < 9; j++) L = fma(D[j], cd[j], L);
                P = P * L;
                if ((kk & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
                    F = F + sleef_log_pos(P);
                    P = 1.0;
                }
                const float r = __builtin_amdgcn_rcpf((float)L);
                const double rd = (double)r;
#pragma unroll
                for (int j = 0; j < 9; j++) G[j] = fma(cd[j], rd, G[j]);
#pragma unroll
                for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
                int id = 0;
#pragma unroll
                for (int a2 = 0; a2 < 9; a2++)
#pragma unroll
                    for (int bb = a2; bb < 9; bb++, id++) h[id] = fmaf(cv[a2], cv[bb], h[id]);
            }
        }
        __syncthreads();
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = G[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = (double)h[k];
        out[(size_t)55 * nslots] = N;
    }
}

// ------------------------------------------------ tri/tetra-allelic sites

// iis.rs:26-104: identity mode of the alleles (i, j) of x and (k, l) of y.
__host__ __device__ constexpr int iis_mode(int i, int j, int k, int l) {
    const int key = ((i == j) << 5) | ((i == k) << 4) | ((i == l) << 3) | ((j == k) << 2) | ((j == l) << 1) | (k == l);
    return key == 0x3f ? 1 : key == 0x21 ? 2 : key == 0x34 ? 3 : key == 0x2a ? 4 : key == 0x20 ? 5 : key == 0x19 ? 6
         : key == 0x07 ? 7 : key == 0x01 ? 8 : key == 0x12 ? 9 : key == 0x0c ? 10 : key == 0x10 ? 11
         : key == 0x08 ? 12 : key == 0x04 ? 13 : key == 0x02 ? 14 : key == 0x00 ? 15 : 0;
}

// iis.rs:106-277 accumulate_iis1..15 split in two: the coefficients of a mode
// depend only on the site's frequencies (iis_coefs writes them once per site,
// in accumulate_iis's expressions and order), and a lane applies them to its
// g = lx ly with the same fma sequence (iis_nnz: how many a mode has;
// iis_target: the component of the 9-vector each one goes to).
__host__ __device__ constexpr int iis_nnz(int mode) {
    return mode == 1 ? 9 : (mode >= 2 && mode <= 4) || mode == 6 || mode == 7 ? 4 : mode == 5 || mode == 8 ? 2
         : mode == 9 || mode == 10 ? 3 : mode >= 11 && mode <= 14 ? 2 : 1;
}

__host__ __device__ constexpr int iis_target(int mode, int t) {
    return mode == 1 ? t
         : mode == 2 ? (t == 0 ? 1 : t == 1 ? 3 : t == 2 ? 5 : 8)
         : mode == 3 || mode == 4 ? (t == 0 ? 2 : t == 1 ? 3 : t == 2 ? 7 : 8)
         : mode == 5 ? (t == 0 ? 3 : 8)
         : mode == 6 || mode == 7 ? (t == 0 ? 4 : t == 1 ? 5 : t == 2 ? 7 : 8)
         : mode == 8 ? (t == 0 ? 5 : 8)
         : mode == 9 || mode == 10 ? 6 + t
         : mode >= 11 && mode <= 14 ? 7 + t : 8;
}

__device__ __forceinline__ void iis_coefs(int mode, double pi, double pj, double pk, double pl, double* o) {
    switch (mode) {
    case 1: {
        const double pi2 = pi * pi, pi3 = pi2 * pi;
        o[0] = pi; o[1] = pi2; o[2] = pi2; o[3] = pi3; o[4] = pi2; o[5] = pi3; o[6] = pi2; o[7] = pi3;
        o[8] = pi2 * pi2;
    } break;
    case 2: {
        const double pipk = pi * pk;
        o[0] = pipk; o[1] = pipk * pk; o[2] = pi * pipk; o[3] = pipk * pipk;
    } break;
    case 3: {
        const double pipl = pi * pl, pi2pl = pi * pipl, two = pi2pl + pi2pl;
        o[0] = pipl; o[1] = two; o[2] = pi2pl; o[3] = two * pi;
    } break;
    case 4: {
        const double pipk = pi * pk, pi2pk = pi * pipk, two = pi2pk + pi2pk;
        o[0] = pipk; o[1] = two; o[2] = pi2pk; o[3] = two * pi;
    } break;
    case 5: {
        const double v = (pi * pk) * pl, two = v + v;
        o[0] = two; o[1] = two * pi;
    } break;
    case 6: {
        const double pipj = pi * pj, pi2pj = pi * pipj, two = pi2pj + pi2pj;
        o[0] = pipj; o[1] = two; o[2] = pi2pj; o[3] = two * pi;
    } break;
    case 7: {
        const double pipj = pi * pj, pj2pi = pj * pipj, two = pj2pi + pj2pi;
        o[0] = pipj; o[1] = two; o[2] = pj2pi; o[3] = two * pj;
    } break;
    case 8: {
        const double v = (pi * pj) * pk, two = v + v;
        o[0] = two; o[1] = two * pk;
    } break;
    case 9:
    case 10: {  // mode 10 is mode 9 (iis.rs:213-215)
        const double pipj = pi * pj, two = pipj + pipj;
        o[0] = two; o[1] = pipj * (pi + pj); o[2] = two * two;
    } break;
    case 11: {
        const double v = (pi * pj) * pl;
        o[0] = v; o[1] = (4.0 * v) * pi;
    } break;
    case 12: {
        const double v = (pi * pj) * pk;
        o[0] = v; o[1] = (4.0 * v) * pi;
    } break;
    case 13: {
        const double v = (pi * pj) * pl;
        o[0] = v; o[1] = (4.0 * v) * pj;
    } break;
    case 14: {
        const double v = (pi * pj) * pk;
        o[0] = v; o[1] = (4.0 * v) * pj;
    } break;
    default: {  // 15
        const double v = (pi * pj) * (pk * pl);
        o[0] = 4.0 * v;
    } break;
    }
}

// The genotype pairs of IIS_A in mixture.rs's order: their IBS modes and the
// offsets of their coefficients in a site's table (NC coefficients a site).
template <int A>
struct IisTable {
    static constexpr int G = A * (A + 1) / 2;
    int mode[G * G];
    int off[G * G + 1];
    constexpr IisTable() : mode(), off() {
        constexpr Genotypes<A> gt;
        off[0] = 0;
        for (int c = 0; c < G * G; c++) {
            const int gx = c / G, gy = c % G;
            mode[c] = iis_mode(gt.gi[gx], gt.gj[gx], gt.gi[gy], gt.gj[gy]);
            off[c + 1] = off[c] + iis_nnz(mode[c]);
        }
    }
};

// One genotype pair's fmas with every index a template constant (a fold over
// the mode's coefficients, left to right: iis.rs's accumulate_iis order).
// (V = double; V = float in the mixed-precision form)
template <int MODE, int OFF, typename V, int... Ts>
__device__ __forceinline__ void iis_apply(V* cv, V g, const V* T, std::integer_sequence<int, Ts...>) {
    ((cv[iis_target(MODE, Ts)] = fma(g, T[OFF + Ts], cv[iis_target(MODE, Ts)])), ...);
}

template <int A, int C, typename V>
__device__ __forceinline__ void iis_pair(const V* lx, const V* ly, const V* T, V* cv) {
    constexpr IisTable<A> tab{};
    constexpr int G = IisTable<A>::G, MODE = tab.mode[C], OFF = tab.off[C];
    const V g = lx[C / G] * ly[C % G];
    iis_apply<MODE, OFF>(cv, g, T, std::make_integer_sequence<int, iis_nnz(MODE)>{});
}

// mixture.rs:8-60 from a site's coefficient table: every (gx, gy) in genotype
// order, g = lx[gx] ly[gy], then its mode's fmas.
template <int A, typename V, int... Cs>
__device__ __forceinline__ void mixture_table(const V* lx, const V* ly, const V* T, V* cv,
                                              std::integer_sequence<int, Cs...>) {
    (iis_pair<A, Cs>(lx, ly, T, cv), ...);
}

// mixture.rs:8-60 with each genotype pair's coefficients computed in
// registers from the site's allele frequencies p[A] by iis_coefs -- the same
// operations as iis_table_kernel, so the same bits as the table's -- instead
// of read from a table (products shared between pairs are computed once).
template <int A, int C>
__device__ __forceinline__ void iis_pair_reg(const double* lx, const double* ly, const double* p, double* cv) {
    constexpr IisTable<A> tab{};
    constexpr Genotypes<A> gt{};
    constexpr int G = IisTable<A>::G, MODE = tab.mode[C], GX = C / G, GY = C % G;
    double o[9];
    iis_coefs(MODE, p[gt.gi[GX]], p[gt.gj[GX]], p[gt.gi[GY]], p[gt.gj[GY]], o);
    const double g = lx[GX] * ly[GY];
    iis_apply<MODE, 0>(cv, g, static_cast<const double*>(o), std::make_integer_sequence<int, iis_nnz(MODE)>{});
}

template <int A, int... Cs>
__device__ __forceinline__ void mixture_reg(const double* lx, const double* ly, const double* p, double* cv,
                                            std::integer_sequence<int, Cs...>) {
    (iis_pair_reg<A, Cs>(lx, ly, p, cv), ...);
}

// Each site's coefficient table: thread (site, genotype pair).
template <int A>
__global__ __launch_bounds__(64) void iis_table_kernel(const double* __restrict__ af, uint32_t n_sites,
                                                        double* __restrict__ table) {
    constexpr IisTable<A> tab;
    constexpr Genotypes<A> gt;
    constexpr int G = IisTable<A>::G, NC = tab.off[G * G];
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n_sites * G * G) return;
    const uint64_t s = e / (G * G);
    const int c = (int)(e % (G * G)), gx = c / G, gy = c % G;
    const double* p = af + s * A;
    double o[9];
    iis_coefs(tab.mode[c], p[gt.gi[gx]], p[gt.gj[gx]], p[gt.gi[gy]], p[gt.gj[gy]], o);
    double* dst = table + s * NC + tab.off[c];
    for (int t = 0; t < iis_nnz(tab.mode[c]); t++) dst[t] = o[t];
}

// The sweep over sites with A = 3 or 4 alleles: as sweep_gl with rows of
// G = A(A+1)/2 likelihoods per site and the generic mixture.  A site's
// coefficient table (NC doubles: 132 for A = 3, 300 for A = 4) is the same for
// every lane; the block stages a chunk's tables in LDS.  Each lane reads its
// own pair's two rows, 8G bytes a site, and loads the next site's rows while
// this site is computed: both rows with A = 3; with A = 4 only y's (both would
// take 40 more VGPRs and spill), as pairs are claimed in sample tiles, so a
// block's lanes share a few x rows, which stay in cache, and read many y rows.
// REG: each genotype pair's coefficients computed in registers from the
// site's allele frequencies (mixture_reg; `table` is then the part's
// frequencies, A a site), the chunk's frequencies staged in LDS.
template <int A, bool REG = false>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_glm(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                           const double* __restrict__ gl, uint32_t v_pad,
                                                           const double* __restrict__ table,
                                                           const uint32_t* __restrict__ seg_bounds,
                                                           const uint32_t* __restrict__ sx,
                                                           const uint32_t* __restrict__ sy,
                                                           const double* __restrict__ pt,
                                                           const int* __restrict__ active, uint32_t nslots,
                                                           double eps, double* __restrict__ partial) {
    constexpr IisTable<A> tab;
    constexpr int G = A * (A + 1) / 2, NC = tab.off[G * G];
    constexpr bool PREFETCH_X = A == 3 && !REG;
    constexpr bool PREFETCH_Y = !REG;
    constexpr int KW = REG ? A : NC;  // doubles a site in LDS
    __shared__ __attribute__((aligned(16))) double kt[CHUNK * KW];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const double* glx = gl + (size_t)x * v_pad * G;
    const double* gly = gl + (size_t)y * v_pad * G;

    double F = 0.0, N = 0.0, P = 1.0, Gr[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) Gr[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;

    // site t's row of one sample (G doubles, 16-byte aligned: G is even)
    auto load_row = [&](const double* base, uint32_t t, double* r) {
        const double2* a = reinterpret_cast<const double2*>(base + (size_t)t * G);
#pragma unroll
        for (int q = 0; q < G / 2; q++) {
            const double2 u = a[q];
            r[2 * q] = u.x;
            r[2 * q + 1] = u.y;
        }
    };
    // segment bounds are whole chunks: t - s0 counts chunks from 0
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    double nx[PREFETCH_X ? G : 1], ny[PREFETCH_Y ? G : 1];
    if (s0 < s1) {
        if (PREFETCH_X) load_row(glx, s0, nx);
        if (PREFETCH_Y) load_row(gly, s0, ny);
    }
    uint2 wx = make_uint2(0, 0), wy = make_uint2(0, 0);
#pragma unroll 1
    for (uint32_t t = s0; t < s1; t++) {
        const int k = (int)((t - s0) & (CHUNK - 1));
        if (k == 0) {
            __syncthreads();  // the previous chunk's tables are read
            {
                const double2* src = reinterpret_cast<const double2*>(table + (size_t)t * KW);
                double2* dst = reinterpret_cast<double2*>(kt);
                for (uint32_t i = tid; i < CHUNK * KW / 2; i += SWEEP_TPB) dst[i] = src[i];
            }
            wx = *reinterpret_cast<const uint2*>(rowx + t / 16);
            wy = *reinterpret_cast<const uint2*>(rowy + t / 16);
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
            __syncthreads();
        }
        const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
        const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
        const bool missing = (ax & 3u) == 3u || (ay & 3u) == 3u;
        double lx[G], ly[G];
#pragma unroll
        for (int q = 0; q < G; q++) {
            if (PREFETCH_X) lx[q] = nx[q];
            if (PREFETCH_Y) ly[q] = ny[q];
        }
        if (!PREFETCH_X) load_row(glx, t, lx);
        if (!PREFETCH_Y) load_row(gly, t, ly);
        if (t + 1 < s1) {
            if (PREFETCH_X) load_row(glx, t + 1, nx);
            if (PREFETCH_Y) load_row(gly, t + 1, ny);
        }
        // mixture.rs:8-60: for every (gx, gy) in genotype order, g = lx ly and
        // accumulate_iis with the site's coefficients of that pair's mode
        double cv[9];
#pragma unroll
        for (int j = 0; j < 9; j++) cv[j] = 0.0;
        if (REG)
            mixture_reg<A>(lx, ly, kt + k * KW, cv, std::make_integer_sequence<int, G * G>{});
        else
            mixture_table<A>(lx, ly, kt + k * KW, cv, std::make_integer_sequence<int, G * G>{});
        site_update(cv, missing ? 1.0 : eps, D, k, P, F, Gr, H);
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = Gr[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
        out[(size_t)55 * nslots] = N;
    }
}

// A tri/tetra-allelic part whose every called row is one-hot (hard calls
// fed as likelihoods: one exact 1.0, the other entries 0) is also held as
// 4-bit genotype indices (8 sites a word, sample-major; 15 = missing): thread
// (sample, word) writes one word and sets *bad when a called row is not
// one-hot.
template <int A>
__global__ __launch_bounds__(256) void part_codes_kernel(const double* __restrict__ rows,
                                                         const uint32_t* __restrict__ mask, uint32_t row_words,
                                                         uint32_t v_pad, uint32_t n_samples,
                                                         uint32_t* __restrict__ codes, int* __restrict__ bad) {
    constexpr int G = A * (A + 1) / 2;
    const uint32_t words = v_pad / 8;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n_samples * words) return;
    const uint32_t smp = (uint32_t)(e / words), w = (uint32_t)(e % words);
    uint32_t out = 0;
    bool ok = true;
    for (int i = 0; i < 8; i++) {
        const uint32_t t = w * 8 + i;
        const uint32_t m = (mask[(size_t)smp * row_words + t / 16] >> (2 * (t % 16))) & 3u;
        uint32_t code = 15u;
        if (m != 3u) {
            const double* r = rows + ((size_t)smp * v_pad + t) * G;
            int ones = 0, zeros = 0, at = 0;
            for (int g = 0; g < G; g++) {
                ones += r[g] == 1.0;
                zeros += r[g] == 0.0;
                if (r[g] == 1.0) at = g;
            }
            ok = ok && ones == 1 && zeros == G - 1;
            code = (uint32_t)at;
        }
        out |= code << (4 * i);
    }
    codes[(size_t)smp * words + w] = out;
    if (!ok) atomicOr(bad, 1);
}

// A hard-call part's rows of c_s for every genotype pair: c_s of a one-hot
// row pair is its genotype pair's coefficients alone -- in mixture.rs:8-60
// every other pair adds g = 0 times its coefficients, exact zeros, and the
// pair's own g is 1.0 -- and each coefficient of a mode goes to a different
// component, so c_s[iis_target(mode, e)] = fma(1, T_e, 0).  Per site G^2 rows
// in genotype order, then a zero row for a missing call; GLC_RW doubles a row
// (9 used, 16-byte aligned).  Thread (site, row).
constexpr int GLC_RW = 10;
template <int A>
__global__ __launch_bounds__(256) void part_lut_kernel(const double* __restrict__ table, uint32_t v_pad,
                                                       double* __restrict__ lut) {
    constexpr IisTable<A> tab;
    constexpr int G = A * (A + 1) / 2, NC = tab.off[G * G], R = G * G + 1;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)v_pad * R) return;
    const uint64_t site = e / R;
    const int c = (int)(e % R);
    double cv[GLC_RW];
#pragma unroll
    for (int q = 0; q < GLC_RW; q++) cv[q] = 0.0;
    if (c < G * G) {
        const int mode = tab.mode[c];
        const double* T = table + site * NC + tab.off[c];
        for (int q = 0; q < iis_nnz(mode); q++) cv[iis_target(mode, q)] = fma(1.0, T[q], 0.0);
    }
    double* dst = lut + e * GLC_RW;
#pragma unroll
    for (int q = 0; q < GLC_RW; q++) dst[q] = cv[q];
}

// sweep_glm for a part held as 4-bit genotype indices (part_codes_kernel):
// a lane reads its genotype pair's row of c_s (part_lut_kernel) instead of
// 2 G likelihoods and the mixture; the site update and the summation order
// are sweep_glm's, so the records are the same bit for bit.  The block stages
// the rows of SUB sites at a time in LDS, the next SUB sites' rows loaded into
// registers while these sites are computed (two LDS buffers); TPB pairs a
// block: 512, one block a CU, so that a CU stages each site's rows once, or
// 256 when blocks of 512 would leave CUs without one (few segments).
constexpr int GLC_TPB = 512;
template <int A, int TPB>
__global__ __launch_bounds__(TPB, 2) void sweep_glm_codes(
    const uint32_t* __restrict__ mask, uint32_t row_words, const uint32_t* __restrict__ codes,
    const double* __restrict__ lutg, const uint32_t* __restrict__ seg_bounds, const uint32_t* __restrict__ sx,
    const uint32_t* __restrict__ sy, const double* __restrict__ pt, const int* __restrict__ active,
    uint32_t nslots, uint32_t v_pad, double eps, double* __restrict__ partial) {
    constexpr int G = A * (A + 1) / 2, R = G * G + 1;  // rows a site
    constexpr int SUB = A == 3 ? 8 : 4;                // sites a stage
    constexpr int N2 = SUB * R * GLC_RW / 2;           // double2 a stage
    constexpr int PER = (N2 + TPB - 1) / TPB;
    static_assert(8 % SUB == 0 && CHUNK % 8 == 0 && (SUB * R * GLC_RW) % 2 == 0,
                  "stages inside a chunk and a code word, whole double2");
    __shared__ __attribute__((aligned(16))) double2 lut[2][N2];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = mask + (size_t)x * row_words;
    const uint32_t* rowy = mask + (size_t)y * row_words;
    const uint32_t* cx = codes + (size_t)x * (v_pad / 8);
    const uint32_t* cy = codes + (size_t)y * (v_pad / 8);

    double F = 0.0, N = 0.0, P = 1.0, Gr[9], H[45];
#pragma unroll
    for (int k = 0; k < 9; k++) Gr[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = 0.0;

    // segment bounds are whole chunks (so whole stages): t - s0 counts chunks from 0
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    double2 nx[PER];
    auto fetch = [&](uint32_t sb) {  // the rows of sites sb .. sb + SUB - 1
        const double2* src = reinterpret_cast<const double2*>(lutg + (size_t)sb * R * GLC_RW);
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t e = tid + (uint32_t)i * TPB;
            if (e < (uint32_t)N2) nx[i] = src[e];
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t e = tid + (uint32_t)i * TPB;
            if (e < (uint32_t)N2) lut[b][e] = nx[i];
        }
    };
    if (s0 < s1) {
        fetch(s0);
        store(0);
    }
    __syncthreads();
    // the mask words of a chunk and the code words of 8 sites, each loaded
    // one chunk / one word ahead of its use (n*: the next)
    uint2 wx = make_uint2(0, 0), wy = make_uint2(0, 0), nwx = wx, nwy = wx;
    uint32_t qx = 0, qy = 0, nqx = 0, nqy = 0;
    if (s0 < s1) {
        nwx = *reinterpret_cast<const uint2*>(rowx + s0 / 16);
        nwy = *reinterpret_cast<const uint2*>(rowy + s0 / 16);
        nqx = cx[s0 / 8];
        nqy = cy[s0 / 8];
    }
    int b = 0;
#pragma unroll 1
    for (uint32_t sb = s0; sb < s1; sb += SUB) {
        if (sb + SUB < s1) fetch(sb + SUB);
        const double* lb = reinterpret_cast<const double*>(lut[b]);
        // a stage lies inside one chunk and one code word (SUB divides 8, and
        // stages start at chunk boundaries): the words move at its first site
        const int k0 = (int)((sb - s0) & (CHUNK - 1));
        if (k0 == 0) {
            wx = nwx;
            wy = nwy;
            if (sb + CHUNK < s1) {
                nwx = *reinterpret_cast<const uint2*>(rowx + (sb + CHUNK) / 16);
                nwy = *reinterpret_cast<const uint2*>(rowy + (sb + CHUNK) / 16);
            }
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
        }
        if (k0 % 8 == 0) {
            qx = nqx;
            qy = nqy;
            if (sb + 8 < s1) {
                nqx = cx[(sb + 8) / 8];
                nqy = cy[(sb + 8) / 8];
            }
        }
        // two sites a pass, so that a site's row is read from LDS while the
        // previous site is computed
#pragma unroll 2
        for (int j = 0; j < SUB; j++) {
            const uint32_t t = sb + (uint32_t)j;
            const int k = k0 + j;
            const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
            const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
            const uint32_t gx = (qx >> (4 * (k & 7))) & 15u, gy = (qy >> (4 * (k & 7))) & 15u;
            const bool missing = (ax & 3u) == 3u || (ay & 3u) == 3u || gx >= (uint32_t)G || gy >= (uint32_t)G;
            const uint32_t row = missing ? (uint32_t)(G * G) : gx * G + gy;
            const double2* src = reinterpret_cast<const double2*>(lb + ((size_t)(t - sb) * R + row) * GLC_RW);
            double cv[9];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double2 u = src[q];
                cv[2 * q] = u.x;
                cv[2 * q + 1] = u.y;
            }
            cv[8] = src[4].x;
            site_update(cv, missing ? 1.0 : eps, D, k, P, F, Gr, H);
        }
        if (sb + SUB < s1) store(b ^ 1);  // its readers finished at the previous barrier
        __syncthreads();
        b ^= 1;
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = Gr[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
        out[(size_t)55 * nslots] = N;
    }
}

// The mixed-precision form of sweep_glm (precision KESTREL_MIXED): the
// coefficient table rounded to FP32 (half the LDS bytes a site: 528 for A =
// 3, 1,200 for A = 4), the likelihoods rounded to FP32 and the generic
// mixture in FP32; then the mixed site update of sweep_gl_mixed (L, the
// gradient terms, the log product and the objective in FP64; 1/L, c' and the
// Hessian summed over the segment in FP32).  Both rows of the next site are
// loaded while this one is computed (FP32 rows leave the registers for it).
template <int A>
__global__ __launch_bounds__(SWEEP_TPB, 2) void sweep_glm_mixed(const uint32_t* __restrict__ geno,
                                                                 uint32_t row_words, const double* __restrict__ gl,
                                                                 uint32_t v_pad, const float* __restrict__ table,
                                                                 const uint32_t* __restrict__ seg_bounds,
                                                                 const uint32_t* __restrict__ sx,
                                                                 const uint32_t* __restrict__ sy,
                                                                 const double* __restrict__ pt,
                                                                 const int* __restrict__ active, uint32_t nslots,
                                                                 double eps, double* __restrict__ partial) {
    constexpr IisTable<A> tab;
    constexpr int G = A * (A + 1) / 2, NC = tab.off[G * G];
    static_assert(NC % 4 == 0, "16-byte table rows");
    __shared__ __attribute__((aligned(16))) float kt[CHUNK * NC];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x * SWEEP_TPB + tid;
    const uint32_t seg = blockIdx.y;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;

    double D[9];
    uint32_t x = 0, y = 0;
    if (live) {
        x = sx[slot];
        y = sy[slot];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = 0.0;
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    const double* glx = gl + (size_t)x * v_pad * G;
    const double* gly = gl + (size_t)y * v_pad * G;

    double F = 0.0, N = 0.0, P = 1.0, Gr[9];
    float h[45];
#pragma unroll
    for (int k = 0; k < 9; k++) Gr[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 45; k++) h[k] = 0.0f;

    // site t's row of one sample, rounded to FP32
    auto load_row = [&](const double* base, uint32_t t, float* r) {
        const double2* a = reinterpret_cast<const double2*>(base + (size_t)t * G);
#pragma unroll
        for (int q = 0; q < G / 2; q++) {
            const double2 u = a[q];
            r[2 * q] = (float)u.x;
            r[2 * q + 1] = (float)u.y;
        }
    };
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    float nx[G], ny[G];
    if (s0 < s1) {
        load_row(glx, s0, nx);
        load_row(gly, s0, ny);
    }
    uint2 wx = make_uint2(0, 0), wy = make_uint2(0, 0);
#pragma unroll 1
    for (uint32_t t = s0; t < s1; t++) {
        const int k = (int)((t - s0) & (CHUNK - 1));
        if (k == 0) {
            __syncthreads();  // the previous chunk's tables are read
            {
                const float4* src = reinterpret_cast<const float4*>(table + (size_t)t * NC);
                float4* dst = reinterpret_cast<float4*>(kt);
                for (uint32_t i = tid; i < CHUNK * NC / 4; i += SWEEP_TPB) dst[i] = src[i];
            }
            wx = *reinterpret_cast<const uint2*>(rowx + t / 16);
            wy = *reinterpret_cast<const uint2*>(rowy + t / 16);
            const uint32_t ux = ~((wx.x & (wx.x >> 1)) | (wy.x & (wy.x >> 1))) & 0x55555555u;
            const uint32_t uy = ~((wx.y & (wx.y >> 1)) | (wy.y & (wy.y >> 1))) & 0x55555555u;
            N += (double)(__popc(ux) + __popc(uy));
            __syncthreads();
        }
        const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
        const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
        const bool missing = (ax & 3u) == 3u || (ay & 3u) == 3u;
        float lx[G], ly[G];
#pragma unroll
        for (int q = 0; q < G; q++) {
            lx[q] = nx[q];
            ly[q] = ny[q];
        }
        if (t + 1 < s1) {
            load_row(glx, t + 1, nx);
            load_row(gly, t + 1, ny);
        }
        float cv[9];
#pragma unroll
        for (int j = 0; j < 9; j++) cv[j] = 0.0f;
        mixture_table<A>(lx, ly, kt + k * NC, cv, std::make_integer_sequence<int, G * G>{});
        double cd[9];
#pragma unroll
        for (int j = 0; j < 9; j++) cd[j] = (double)cv[j];
        double L = missing ? 1.0 : eps;
#pragma unroll
        for (int j = 0; j < 9; j++) L = fma(D[j], cd[j], L);
        P = P * L;
        if ((k & (LOG_GROUP - 1)) == LOG_GROUP - 1) {
            F = F + sleef_log_pos(P);
            P = 1.0;
        }
        const float r = __builtin_amdgcn_rcpf((float)L);
        const double rd = (double)r;
#pragma unroll
        for (int j = 0; j < 9; j++) Gr[j] = fma(cd[j], rd, Gr[j]);
#pragma unroll
        for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
        int id = 0;
#pragma unroll
        for (int a2 = 0; a2 < 9; a2++)
#pragma unroll
            for (int bb = a2; bb < 9; bb++, id++) h[id] = fmaf(cv[a2], cv[bb], h[id]);
    }
    if (live) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
#pragma unroll
        for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = Gr[k];
#pragma unroll
        for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = (double)h[k];
        out[(size_t)55 * nslots] = N;
    }
}

// FP32 copy of a table (the mixed form's coefficients)
__global__ __launch_bounds__(256) void round_f32_kernel(const double* __restrict__ src, float* __restrict__ dst,
                                                        uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

// Evaluation at points outside the sweeps' domain (any point, any eps:
// L = eps + c.D may be zero, negative, subnormal or not finite), in the
// sweeps' segmented order with the reference's arithmetic at every site:
// 1/L by IEEE division and, where a 16-site group holds an L that is not a
// positive normal number or its product is not one, the group's logs taken
// one site at a time with the full log (0 -> -inf, negative or NaN -> NaN,
// log/scalar.rs:49-57), added in site order; otherwise the log of the
// group's product as in the sweeps.  On the sweeps' domain it returns their
// FP64 bits.  One thread per (slot, segment); slots with exact[slot] == 0
// are left to the sweeps.  KIND: 0 = 2-bit rows (c from the site's table
// block), 2 / 3 = likelihood rows of stride 2 / 3 (mixture_row2), 13 / 14 =
// a part of 3- / 4-allele sites (the generic mixture over its tables).
template <int KIND>
__global__ __launch_bounds__(64) void sweep_exact_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                   const double* __restrict__ gl, uint32_t v_pad,
                                                   const double* __restrict__ af_or_table,
                                                   const uint32_t* __restrict__ seg_bounds,
                                                   const uint32_t* __restrict__ sx, const uint32_t* __restrict__ sy,
                                                   const double* __restrict__ pt, const int* __restrict__ exact,
                                                   uint32_t nslots, double eps, double* __restrict__ partial) {
    constexpr int A = KIND == 13 ? 3 : KIND == 14 ? 4 : 2;
    constexpr int G = A * (A + 1) / 2;
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
    const uint32_t seg = blockIdx.y;
    if (slot >= nslots || !exact[slot]) return;
    const uint32_t x = sx[slot], y = sy[slot];
    double D[9];
    for (int k = 0; k < 9; k++) D[k] = pt[(size_t)k * nslots + slot];
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    double F = 0.0, N = 0.0, P = 1.0, Gr[9], H[45], Lg[LOG_GROUP];
    bool normal = true;
    for (int k = 0; k < 9; k++) Gr[k] = 0.0;
    for (int k = 0; k < 45; k++) H[k] = 0.0;
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    for (uint32_t t = s0; t < s1; t++) {
        const uint32_t cx = (rowx[t / 16] >> (2 * (t % 16))) & 3u, cy = (rowy[t / 16] >> (2 * (t % 16))) & 3u;
        const bool missing = cx == 3u || cy == 3u;
        double cv[9];
        for (int j = 0; j < 9; j++) cv[j] = 0.0;
        if (KIND == 0) {
            if (!missing) lut_block(af_or_table[2 * t], af_or_table[2 * t + 1], (int)(cx * 3 + cy), cv);
        } else {
            constexpr int ST = KIND == 2 ? 2 : G;
            const double* px = gl + ((size_t)x * v_pad + t) * ST;
            const double* py = gl + ((size_t)y * v_pad + t) * ST;
            double lx[G], ly[G];
            if (KIND == 2) {  // the exact 1.0 back at the position of the 2-bit code
                lx[0] = cx == 0 ? 1.0 : px[0];
                lx[1] = cx == 0 ? px[0] : cx == 1 ? 1.0 : px[1];
                lx[2] = cx == 2 ? 1.0 : px[1];
                ly[0] = cy == 0 ? 1.0 : py[0];
                ly[1] = cy == 0 ? py[0] : cy == 1 ? 1.0 : py[1];
                ly[2] = cy == 2 ? 1.0 : py[1];
            } else {
                for (int q = 0; q < G; q++) {
                    lx[q] = px[q];
                    ly[q] = py[q];
                }
            }
            if (KIND == 2 || KIND == 3) {
                double K[K_N];
                site_coefficients(af_or_table[2 * t], af_or_table[2 * t + 1], K);
                mixture_soft(lx, ly, K, cv);
            } else {
                constexpr IisTable<A> tab{};
                mixture_table<A>(lx, ly, af_or_table + (size_t)t * tab.off[G * G], cv,
                                 std::make_integer_sequence<int, G * G>{});
            }
        }
        // a site with a missing call is not a row (lanevector.rs:100-104):
        // it adds nothing (the sweeps' L = 1, c = 0 there, which is the same
        // on their domain)
        const int k = (int)((t - s0) & (LOG_GROUP - 1));
        Lg[k] = 1.0;
        if (!missing) {
            N += 1.0;
            double L = eps;
            for (int j = 0; j < 9; j++) L = fma(D[j], cv[j], L);
            P = P * L;
            Lg[k] = L;
            normal = normal && L >= 2.2250738585072014e-308 && L <= 1.7976931348623157e308;
            const double r = 1.0 / L;
            for (int j = 0; j < 9; j++) Gr[j] = fma(cv[j], r, Gr[j]);
            for (int j = 0; j < 9; j++) cv[j] = cv[j] * r;
            int id = 0;
            for (int a = 0; a < 9; a++)
                for (int b = a; b < 9; b++, id++) H[id] = fma(cv[a], cv[b], H[id]);
        }
        if (k == LOG_GROUP - 1) {
            if (normal && P >= 2.2250738585072014e-308 && P <= 1.7976931348623157e308) {
                F = F + sleef_log_pos(P);
            } else {
                for (int q = 0; q < LOG_GROUP; q++) F = F + sleef_log(Lg[q]);
            }
            P = 1.0;
            normal = true;
        }
    }
    double* out = partial + (size_t)seg * NPART * nslots + slot;
    out[0] = F;
    for (int k = 0; k < 9; k++) out[(size_t)(1 + k) * nslots] = Gr[k];
    for (int k = 0; k < 45; k++) out[(size_t)(10 + k) * nslots] = H[k];
    out[(size_t)55 * nslots] = N;
}

// A compact 16-byte row: the two entries other than the first exact 1.0, in
// order; the 1.0's position is the row's 2-bit code (3 = missing: the stored
// pair is 0, 0).  A called row with no exact 1.0 sets *lacks (the caller then
// builds 24-byte rows instead).
__device__ __forceinline__ uint32_t store_compact(double* dst, const double* l, uint32_t code, int* lacks) {
    if (code != 3) {
        code = l[0] == 1.0 ? 0u : l[1] == 1.0 ? 1u : l[2] == 1.0 ? 2u : 3u;
        if (code == 3) {
            *lacks = 1;
            code = 0;
        }
    }
    dst[0] = code == 0 ? l[1] : l[0];
    dst[1] = code == 2 ? l[1] : l[2];
    return code;
}

// Re-bucketed rows (wrapper.rs:237-410 remove_zero_frequencies): new site t of
// a part with Gp genotypes takes, from site map_site[t] of the input bucket
// with map_alleles[t] alleles, the likelihoods at gmap[t][0..Gp) (the
// genotypes of the alleles with non-zero frequency); thread (sample i, word w)
// writes sites 16w..16w+15 of row i (0 for a missing call, a NaN among the
// selected entries) and the 2-bit mask word (3 = missing, padding = 3).
// With compact (3 genotypes only) the rows are 16-byte ones as store_compact
// makes them.
__global__ __launch_bounds__(256) void gather_multi_kernel(const double* __restrict__ src2,
                                                            const double* __restrict__ src3,
                                                            const double* __restrict__ src4, uint32_t n_samples,
                                                            const uint32_t* __restrict__ map_site,
                                                            const uint8_t* __restrict__ map_alleles,
                                                            const uint8_t* __restrict__ gmap, uint32_t n_t, int gp,
                                                            double* __restrict__ rows, uint32_t v_pad,
                                                            uint32_t* __restrict__ mask, uint32_t row_words,
                                                            int compact, int* __restrict__ no_one) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_samples * row_words) return;
    const uint32_t i = (uint32_t)(tid / row_words), w = (uint32_t)(tid % row_words);
    uint32_t word = 0;
    int lacks_one = 0;
    for (int j = 0; j < 16; j++) {
        const uint32_t t = w * 16 + j;
        uint32_t code = 3;
        if (compact) {  // gp == 3: the biallelic part
            double l[3] = {0.0, 0.0, 0.0};
            if (t < n_t) {
                const int na = map_alleles[t];
                const int gs = na * (na + 1) / 2;
                const double* src =
                    (na == 2 ? src2 : na == 3 ? src3 : src4) + ((size_t)map_site[t] * n_samples + i) * gs;
                for (int k = 0; k < 3; k++) l[k] = src[gmap[(size_t)t * 3 + k]];
                if (!(isnan(l[0]) || isnan(l[1]) || isnan(l[2])))
                    code = 0;
                else
                    l[0] = l[1] = l[2] = 0.0;
            }
            code = store_compact(rows + ((size_t)i * v_pad + t) * 2, l, code, &lacks_one);
            word |= code << (2 * j);
            continue;
        }
        double* dst = rows + ((size_t)i * v_pad + t) * gp;
        if (t < n_t) {
            const int na = map_alleles[t];
            const int gs = na * (na + 1) / 2;
            const double* src = (na == 2 ? src2 : na == 3 ? src3 : src4) + ((size_t)map_site[t] * n_samples + i) * gs;
            bool nan = false;
            for (int k = 0; k < gp; k++) nan |= isnan(src[gmap[(size_t)t * gp + k]]);
            if (!nan) {
                code = 0;
                for (int k = 0; k < gp; k++) dst[k] = src[gmap[(size_t)t * gp + k]];
            }
        }
        if (code == 3)
            for (int k = 0; k < gp; k++) dst[k] = 0.0;
        word |= code << (2 * j);
    }
    mask[(size_t)i * row_words + w] = word;
    if (lacks_one) atomicOr(no_one, 1);
}

// Missing calls (code 3) among the first n_sites sites of every row: one
// thread per row word (padding sites excluded), one atomic per wavefront.
__global__ __launch_bounds__(256) void count_missing_kernel(const uint32_t* __restrict__ geno, uint32_t n_samples,
                                                             uint32_t row_words, uint32_t n_sites,
                                                             unsigned long long* __restrict__ count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0;
    if (i < (size_t)n_samples * row_words) {
        const uint32_t first = (uint32_t)(i % row_words) * 16;
        const uint32_t v = geno[i];
        uint32_t m = v & (v >> 1) & 0x55555555u;
        if (first + 16 > n_sites) {
            const uint32_t keep = n_sites > first ? n_sites - first : 0;
            m &= (1u << (2 * keep)) - 1u;  // keep < 16
        }
        c = (uint32_t)__popc(m);
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (__lane_id() == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// Soft rows of the kept sites: thread (sample i, word w) moves sites
// 16w..16w+15 from the site-major input (V x N x 3) into the sample-major
// rows (0 for a missing call or a padding site) and writes the 2-bit mask word
// (3 = a NaN in the triple, lanevector.rs:100-104; padding = 3).
__global__ __launch_bounds__(256) void gather_gl_kernel(const double* __restrict__ gl_site, uint32_t n_samples,
                                                         const uint32_t* __restrict__ kept, uint32_t n_kept,
                                                         double* __restrict__ rows, uint32_t v_pad,
                                                         uint32_t* __restrict__ mask, uint32_t row_words, int compact,
                                                         int* __restrict__ no_one) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_samples * row_words) return;
    const uint32_t i = (uint32_t)(tid / row_words), w = (uint32_t)(tid % row_words);
    uint32_t word = 0;
    int lacks_one = 0;
    for (int j = 0; j < 16; j++) {
        const uint32_t t = w * 16 + j;
        double l[3] = {0.0, 0.0, 0.0};
        uint32_t code = 3;
        if (t < n_kept) {
            const double* src = gl_site + ((size_t)kept[t] * n_samples + i) * 3;
            const double a = src[0], b2 = src[1], c = src[2];
            if (!(isnan(a) || isnan(b2) || isnan(c))) {
                l[0] = a;
                l[1] = b2;
                l[2] = c;
                code = 0;
            }
        }
        if (compact) {
            code = store_compact(rows + ((size_t)i * v_pad + t) * 2, l, code, &lacks_one);
        } else {
            double* dst = rows + ((size_t)i * v_pad + t) * 3;
            dst[0] = l[0];
            dst[1] = l[1];
            dst[2] = l[2];
        }
        word |= code << (2 * j);
    }
    mask[(size_t)i * row_words + w] = word;
    if (lacks_one) atomicOr(no_one, 1);
}

// *nonzero = 1 where any of rows[0..n) is not zero (the 16-byte rows of a
// likelihood cohort hold the two entries beside the exact 1.0: all zero means
// every called row is one-hot)
__global__ __launch_bounds__(256) void any_nonzero_kernel(const double* __restrict__ rows, uint64_t n,
                                                          int* __restrict__ nonzero) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool nz = i < n && rows[i] != 0.0;
    if (__ballot(nz) && __lane_id() == 0) atomicOr(nonzero, 1);
}

// ------------------------------------------------------------ SQP step

namespace {

__device__ __forceinline__ void fold(const double* __restrict__ partial, uint32_t nseg, uint32_t nslots,
                                     uint32_t slot, double* acc /* NPART */) {
    for (int k = 0; k < NPART; k++) acc[k] = partial[(size_t)k * nslots + slot];
    for (uint32_t s = 1; s < nseg; s++) {
        const double* p = partial + (size_t)s * NPART * nslots + slot;
        for (int k = 0; k < NPART; k++) acc[k] = acc[k] + p[(size_t)k * nslots];
    }
}

__device__ __forceinline__ void write_point(double* pt, uint32_t nslots, uint32_t slot, const double* v) {
    for (int k = 0; k < 9; k++) pt[(size_t)k * nslots + slot] = v[k];
}

}  // namespace

// Claims of a workgroup's lanes from runs (the step kernel's workgroup is a
// sweep workgroup's 256 slots): the lanes that want a pair take consecutive
// claim indices, first from what is left of the workgroup's run, then from
// the next unclaimed run (next_run counts runs handed out; a run is at least
// 256 claims long unless it is the last, so one new run covers the rest).
// Every thread of the workgroup calls it.  idx: the lane's claim index, or
// -1; exhausted: no run was left for it.
__device__ __forceinline__ void group_claim(const RunClaim& rc, unsigned long long* __restrict__ next_run, bool want,
                                            int64_t& idx, bool& exhausted) {
    __shared__ unsigned int wcnt[SWEEP_TPB / 64];
    __shared__ unsigned long long grant[5];  // b1, n1, b2, n2, exhausted
    const unsigned long long m = __ballot(want);
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x / 64), nw = (int)(blockDim.x / 64);
    if (lane == 0) wcnt[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long cnt = 0;
        for (int w = 0; w < nw; w++) cnt += wcnt[w];
        unsigned long long* g = rc.wrun + 2 * blockIdx.x;
        unsigned long long cur = g[0], end = g[1], b2 = 0, n2 = 0, ex = 0;
        const unsigned long long n1 = min(cnt, end - cur), b1 = cur;
        cur += n1;
        if (n1 < cnt) {
            const unsigned long long r = atomicAdd(next_run, 1ull);
            if (r < rc.nruns) {
                cur = rc.runs[2 * r];
                end = rc.runs[2 * r + 1];
                b2 = cur;
                n2 = min(cnt - n1, end - cur);
                cur += n2;
            } else {
                ex = 1;
            }
        }
        if (cnt) {
            g[0] = cur;
            g[1] = end;
        }
        grant[0] = b1;
        grant[1] = n1;
        grant[2] = b2;
        grant[3] = n2;
        grant[4] = ex;
    }
    __syncthreads();
    if (want) {
        unsigned long long rank = (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
        for (int w = 0; w < wave; w++) rank += wcnt[w];
        const unsigned long long b1 = grant[0], n1 = grant[1], b2 = grant[2], n2 = grant[3];
        if (rank < n1)
            idx = (int64_t)(b1 + rank);
        else if (rank - n1 < n2)
            idx = (int64_t)(b2 + (rank - n1));
        else
            exhausted = grant[4] != 0;
    }
    __syncthreads();  // wcnt and grant are free for the next call
}

template <bool RUNS>
__device__ __forceinline__ void step_slot(SlotState* __restrict__ st, uint32_t* __restrict__ sx,
                                          uint32_t* __restrict__ sy, double* __restrict__ pt,
                                          int* __restrict__ active, uint32_t nslots,
                                          const double* __restrict__ partial, uint32_t nseg,
                                          const uint32_t* __restrict__ pairs, uint64_t npairs,
                                          const uint64_t* __restrict__ order,
                                          unsigned long long* __restrict__ next_pair,
                                          kestrel_output* __restrict__ out, const Tune& tune,
                                          int* __restrict__ n_live, const PairQueue& q, uint32_t slot,
                                          SlotState& s, const RunClaim& rc, bool alive) {
    // RUNS: batch claims in runs, by the whole workgroup (step_runs_kernel);
    // alive: false for a retired slot (or one past nslots) there, which only
    // takes part in the claim
    bool claim = alive && (s.phase == PH_IDLE || s.phase == PH_WAIT);
    if (alive && !claim) {
        // segment sums, left in segment 0 by fold_kernel
        double acc[NPART];
        for (int k = 0; k < NPART; k++) acc[k] = partial[(size_t)k * nslots + slot];
        double e[9];
        for (int k = 0; k < 9; k++) e[k] = pt[(size_t)k * nslots + slot];
        const double n = acc[55];
        kestrel_output o;
        o.x = sx[slot];
        o.y = sy[slot];
        bool done = false;
        if (n == 0.0) {  // lanevector.rs:133-138 would panic: no usable site
            const double nan = __longlong_as_double(0x7ff8000000000000LL);
            for (int k = 0; k < 9; k++) o.jacquard[k] = nan;
            o.obj = nan;
            o.iters = 0;
            o.cond = nan;
            o.n_sites = 0;
            o.status = KESTREL_PAIR_NO_SITES;
            out[s.pair] = o;
            if (q.cap) {
                atomicAdd(&q.req_done[s.req], 1u);
                atomicSub(q.busy, 1);
            }
            claim = true;
        } else {
            const double fe = -acc[0] / n;
            double g[9], h[9][9];
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = -acc[1 + k] / n;
            {
                int id = 0;
#pragma unroll
                for (int a = 0; a < 9; a++)
#pragma unroll
                    for (int b = a; b < 9; b++, id++) {
                        const double v = acc[10 + id] / n;
                        h[a][b] = v;
                        h[b][a] = v;
                    }
            }
            bool accepted;
            if (s.phase == PH_INIT) {
                for (int k = 0; k < 9; k++) s.x[k] = e[k];
                s.f = fe;
                s.k = 0;
                accepted = true;
            } else if (fe <= s.f + s.alpha * s.t || s.bls + 1 == tune.bls_max_iter) {
                for (int k = 0; k < 9; k++) s.x[k] = e[k];
                s.f = fe;
                s.k += 1;
                accepted = true;
            } else {
                s.bls += 1;
                s.alpha *= tune.bls_step_size_reduce;
                double xn[9];
                for (int k = 0; k < 9; k++) xn[k] = s.x[k] + s.alpha * s.p[k];
                sum_to_one<9>(xn, xn);
                write_point(pt, nslots, slot, xn);
                accepted = false;
            }
            if (accepted) {
                double xr[9];
#pragma unroll
                for (int k = 0; k < 9; k++) xr[k] = s.x[k];
                if (s.k == tune.sqp_max_iter) {
                    o.iters = tune.sqp_max_iter;
                    o.cond = __longlong_as_double(0x7ff8000000000000LL);
                    done = true;
                } else if (converged<9>(xr, g, tune.sqp_conv_tol)) {
                    o.iters = s.k;
                    o.cond = condition_reg<9>(xr, g, h, tune.sqp_conv_tol);
                    if (tune.recond_margin > 0.0 &&
                        complementarity_marginal<9>(xr, g, tune.sqp_conv_tol, tune.recond_margin))
                        o.cond = -o.cond;
                    done = true;
                } else {
                    double c[9], yq[9];
#pragma unroll
                    for (int i = 0; i < 9; i++) {
                        double hx = 0.0;
#pragma unroll
                        for (int j = 0; j < 9; j++) hx = fma(h[i][j], xr[j], hx);
                        c[i] = g[i] - hx;
                    }
                    solve_qp_reg<9>(h, c, xr, tune, yq);
                    double gp = 0.0;
#pragma unroll
                    for (int i = 0; i < 9; i++) {
                        s.p[i] = yq[i] - xr[i];
                        gp = fma(g[i], s.p[i], gp);
                    }
                    s.t = tune.bls_sufficient_decrease * gp;
                    s.alpha = 1.0;
                    s.bls = 0;
                    double xn[9];
                    for (int k = 0; k < 9; k++) xn[k] = s.x[k] + s.alpha * s.p[k];
                    sum_to_one<9>(xn, xn);
                    write_point(pt, nslots, slot, xn);
                    s.phase = PH_TRIAL;
                }
                if (done) {
                    for (int k = 0; k < 9; k++) o.jacquard[k] = s.x[k];
                    o.obj = s.f;
                    o.n_sites = (uint32_t)n;
                    o.status = KESTREL_PAIR_OK;
                    out[s.pair] = o;
                    if (q.cap) {
                        atomicAdd(&q.req_done[s.req], 1u);
                        atomicSub(q.busy, 1);
                    }
                    claim = true;
                }
            }
        }
    }
    // batch claims in runs: every thread of the workgroup reaches this point
    int64_t run_idx = -1;
    bool run_out = false;
    if (RUNS) group_claim(rc, next_pair, claim && !q.cap, run_idx, run_out);
    if (!alive) return;
    if (claim) {
        bool got = false;
        uint32_t px = 0, py = 0;
        if (RUNS && !q.cap) {
            if (run_idx >= 0) {
                const uint64_t p = order ? order[run_idx] : (uint64_t)run_idx;
                s.pair = (int64_t)p;
                px = pairs[2 * p];
                py = pairs[2 * p + 1];
                got = true;
            } else if (run_out) {
                s.phase = PH_RETIRED;
                s.pair = -1;
                atomicSub(n_live, 1);
            } else {
                // the new run was the short last one: claim at the next step
                s.phase = PH_IDLE;
                s.pair = -1;
            }
        } else if (q.cap) {
            const unsigned long long idx =
                s.phase == PH_WAIT ? (unsigned long long)s.pair : atomicAdd(next_pair, 1ull);
            const unsigned long long tail = __atomic_load_n(q.tail, __ATOMIC_RELAXED);
            if (idx < tail) {
                const uint4 e = *reinterpret_cast<const uint4*>(q.ent + (idx % q.cap) * 4);
                px = e.x;
                py = e.y;
                s.pair = (int64_t)e.z;
                s.req = e.w;
                got = true;
                atomicAdd(q.busy, 1);
            } else {
                s.phase = PH_WAIT;
                s.pair = (int64_t)idx;
            }
        } else {
            const unsigned long long idx = atomicAdd(next_pair, 1ull);
            if (idx < npairs) {
                // the idx-th claim fits pair order[idx] (a claim order that
                // keeps the sample rows in flight few); its record goes to its
                // own position
                const uint64_t p = order ? order[idx] : (uint64_t)idx;
                s.pair = (int64_t)p;
                px = pairs[2 * p];
                py = pairs[2 * p + 1];
                got = true;
            } else {
                s.phase = PH_RETIRED;
                s.pair = -1;
                atomicSub(n_live, 1);
            }
        }
        if (got) {
            sx[slot] = px;
            sy[slot] = py;
            // coefficients.rs:88-92 start point, normalised as sqp.rs:54
            double x0[9];
            for (int k = 0; k < 9; k++) x0[k] = px == py ? 0.0 : 1.0 / 9.0;
            if (px == py) {
                x0[0] = 0.5;
                x0[6] = 0.5;
            }
            sum_to_one<9>(x0, x0);
            write_point(pt, nslots, slot, x0);
            s.phase = PH_INIT;
            s.k = 0;
            s.bls = 0;
        }
    }
    active[slot] = (s.phase == PH_INIT || s.phase == PH_TRIAL) ? 1 : 0;
    st[slot] = s;
}

// Sums the segment partials of every swept slot in segment order (the
// summation order of fold()) into segment 0, one thread per (partial, slot).
__global__ __launch_bounds__(256) void fold_kernel(double* __restrict__ partial, uint32_t nseg, uint32_t nslots,
                                                    const int* __restrict__ active) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)NPART * nslots || !active[i % nslots]) return;
    double* p = partial + i;  // i = k * nslots + slot
    double a = p[0];
#pragma unroll 4
    for (uint32_t sg = 1; sg < nseg; sg++) a = a + p[(size_t)sg * NPART * nslots];
    p[0] = a;
}

__global__ __launch_bounds__(64) void step_kernel(SlotState* __restrict__ st, uint32_t* __restrict__ sx,
                                                   uint32_t* __restrict__ sy, double* __restrict__ pt,
                                                   int* __restrict__ active, uint32_t nslots,
                                                   const double* __restrict__ partial, uint32_t nseg,
                                                   const uint32_t* __restrict__ pairs, uint64_t npairs,
                                                   const uint64_t* __restrict__ order,
                                                   unsigned long long* __restrict__ next_pair,
                                                   kestrel_output* __restrict__ out, Tune tune,
                                                   int* __restrict__ n_live, PairQueue q, int* __restrict__ live_out) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    SlotState s = st[slot];
    if (s.phase == PH_RETIRED) return;
    step_slot<false>(st, sx, sy, pt, active, nslots, partial, nseg, pairs, npairs, order, next_pair, out, tune, n_live,
                     q, slot, s, RunClaim{}, true);
    if (live_out) {
        // slots still holding or claiming pairs after this step, one atomic
        // per wave (retired lanes have returned above)
        const unsigned long long act = __ballot(1);
        const unsigned long long m = __ballot(s.phase != PH_RETIRED);
        if (__lane_id() == __ffsll(act) - 1 && m) atomicAdd(live_out, (int)__popcll(m));
    }
}

// The step of a batch fit that claims in runs: a workgroup is a sweep
// workgroup's 256 slots, and no thread leaves before the workgroup's claim.
__global__ __launch_bounds__(SWEEP_TPB) void step_runs_kernel(
    SlotState* __restrict__ st, uint32_t* __restrict__ sx, uint32_t* __restrict__ sy, double* __restrict__ pt,
    int* __restrict__ active, uint32_t nslots, const double* __restrict__ partial, uint32_t nseg,
    const uint32_t* __restrict__ pairs, uint64_t npairs, const uint64_t* __restrict__ order,
    unsigned long long* __restrict__ next_run, kestrel_output* __restrict__ out, Tune tune, int* __restrict__ n_live,
    int* __restrict__ live_out, RunClaim rc) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    SlotState s{};
    const bool in = slot < nslots;
    if (in) s = st[slot];
    const bool alive = in && s.phase != PH_RETIRED;
    step_slot<true>(st, sx, sy, pt, active, nslots, partial, nseg, pairs, npairs, order, next_run, out, tune, n_live,
                    PairQueue{}, slot, s, rc, alive);
    if (live_out) {
        const unsigned long long m = __ballot(alive && s.phase != PH_RETIRED);
        if (__lane_id() == 0 && m) atomicAdd(live_out, (int)__popcll(m));
    }
}

// Explicit evaluation points: fold partials into f, g, H (objective.rs:34-39,
// grad_hess.rs:20-38 normalisation).
__global__ __launch_bounds__(64) void finalize_eval(const double* __restrict__ partial, uint32_t nseg,
                                                     uint32_t nslots, uint32_t npts, double* __restrict__ f,
                                                     double* __restrict__ g, double* __restrict__ h,
                                                     uint32_t* __restrict__ nsites) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double acc[NPART];
    fold(partial, nseg, nslots, i, acc);
    const double n = acc[55];
    f[i] = -acc[0] / n;
    for (int k = 0; k < 9; k++) g[(size_t)i * 9 + k] = -acc[1 + k] / n;
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int b = a; b < 9; b++, id++) {
            const double v = acc[10 + id] / n;
            h[(size_t)i * 81 + a * 9 + b] = v;
            h[(size_t)i * 81 + b * 9 + a] = v;
        }
    nsites[i] = (uint32_t)n;
}

// Condition estimates from FP64 sweep partials at given final points (the
// mixed-precision form's converged pairs whose estimate from the FP32-summed
// Hessian reached its resolution): fold, normalise as the step does
// (grad_hess.rs:20-38), then estimate_condition_number (sqp.rs:312-391) at x.
__global__ __launch_bounds__(64) void recondition_kernel(const double* __restrict__ partial, uint32_t nseg,
                                                          uint32_t nslots, uint32_t npts,
                                                          const double* __restrict__ pt, double tol,
                                                          double* __restrict__ cond) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double acc[NPART];
    fold(partial, nseg, nslots, i, acc);
    const double n = acc[55];
    double x[9], g[9], h[9][9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        x[k] = pt[(size_t)k * nslots + i];
        g[k] = -acc[1 + k] / n;
    }
    int id = 0;
#pragma unroll
    for (int a = 0; a < 9; a++)
#pragma unroll
        for (int b = a; b < 9; b++, id++) {
            const double v = acc[10 + id] / n;
            h[a][b] = v;
            h[b][a] = v;
        }
    cond[i] = condition_reg<9>(x, g, h, tol);
}

// Self-test of the sweep's arithmetic shortcuts: recip_rn against IEEE
// division and sleef_log_pos against the full sleef_log, on n inputs drawn
// from a counter-based hash (log-uniform over [1e-9, 1e4], and [1.75, 2.5),
// the range of 1 + m inside the log).
__global__ __launch_bounds__(256) void selftest_kernel(uint64_t n, uint64_t seed,
                                                        unsigned long long* __restrict__ bad) {
    unsigned long long local = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        const double x = (i & 1) ? exp2(-29.9 + u * 43.2) : 1.75 + 0.75 * u;
        if (__double_as_longlong(recip_rn(x)) != __double_as_longlong(1.0 / x)) local++;
        if (__double_as_longlong(sleef_log_pos(x)) != __double_as_longlong(sleef_log(x))) local++;
    }
    if (local) atomicAdd(bad, local);
}

// Achievable vector FMA throughput: 16 independent fma chains per lane, in
// FP64 or (as this library compiles it: no packed FP32) FP32.
template <class T>
__global__ __launch_bounds__(256) void calib_fma_kernel(uint32_t iters, T a, T b, T* __restrict__ sink) {
    T acc[16];
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = (T)(threadIdx.x + k);
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 16; k++) acc[k] = fma(acc[k], a, b);
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) s += acc[k];
    if (s == (T)1.2345) sink[0] = s;  // keeps the chains live
}

// LDS read bandwidth: each lane reads 16 bytes per load, a wave 1 KB of
// consecutive addresses (no bank conflicts); the 8 loads of a step read rows
// that change every step, so none can be hoisted out of the loop.
__global__ __launch_bounds__(256) void calib_lds_kernel(uint32_t iters, double* __restrict__ sink) {
    __shared__ __attribute__((aligned(16))) double buf[8192];  // 64 KB: 64 rows of 64 double2
    for (uint32_t i = threadIdx.x; i < 8192; i += 256) buf[i] = (double)i;
    __syncthreads();
    const double2* b2 = reinterpret_cast<const double2*>(buf);
    const uint32_t lane = threadIdx.x & 63u;
    double2 acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = make_double2(0.0, 0.0);
    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t row = (it * 13u) & 63u;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double2 v = b2[((row + 8u * k) & 63u) * 64u + lane];
            acc[k].x += v.x;
            acc[k].y += v.y;
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) s += acc[k].x + acc[k].y;
    if (s == 1.2345) sink[0] = s;  // keeps the loads live
}

// Plain copy of n16 16-byte words (the reference kernel for HBM byte counters).
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------ CLS (-f GT)

namespace {

// iis.rs:279-462 conditional_probabilities, every IBS mode, in the
// reference's operation order.
__device__ __forceinline__ void cond_prob(double pi, double pj, double pk, double pl, int mode, double* o) {
    const double pi2 = pi * pi, pi3 = pi2 * pi, pj2 = pj * pj, pk2 = pk * pk;
#pragma unroll
    for (int i = 0; i < 9; i++) o[i] = 0.0;
    switch (mode) {
    case 1:
        o[0] = pi; o[1] = pi2; o[2] = pi2; o[3] = pi3; o[4] = pi2; o[5] = pi3; o[6] = pi2; o[7] = pi3;
        o[8] = pi2 * pi2;
        break;
    case 2: o[1] = pi * pk; o[3] = pi * pk2; o[5] = pi2 * pk; o[8] = pi2 * pk2; break;
    case 3: o[2] = pi * pl; o[3] = (2.0 * pi2) * pl; o[7] = pi2 * pl; o[8] = (2.0 * pi3) * pl; break;
    case 4: o[2] = pi * pk; o[3] = (2.0 * pi2) * pk; o[7] = pi2 * pk; o[8] = (2.0 * pi3) * pk; break;
    case 5: o[3] = ((2.0 * pi) * pk) * pl; o[8] = ((2.0 * pi2) * pk) * pl; break;
    case 6: o[4] = pi * pj; o[5] = (2.0 * pi2) * pj; o[7] = pi2 * pj; o[8] = (2.0 * pi3) * pj; break;
    case 7: o[4] = pi * pj; o[5] = (2.0 * pj2) * pi; o[7] = pj2 * pi; o[8] = (2.0 * (pj2 * pj)) * pi; break;
    case 8: o[5] = ((2.0 * pi) * pj) * pk; o[8] = ((2.0 * pk2) * pi) * pj; break;
    case 9:
    case 10:  // ij-ij and ij-ji
        o[6] = (2.0 * pi) * pj; o[7] = (pi * pj) * (pi + pj); o[8] = (4.0 * pi2) * pj2;
        break;
    case 11: o[7] = (pi * pj) * pl; o[8] = ((4.0 * pi2) * pj) * pl; break;
    case 12: o[7] = (pi * pj) * pk; o[8] = ((4.0 * pi2) * pj) * pk; break;
    case 13: o[7] = (pi * pj) * pl; o[8] = ((4.0 * pi) * pj2) * pl; break;
    case 14: o[7] = (pi * pj) * pk; o[8] = ((4.0 * pi) * pj2) * pk; break;
    default:  // 15: ij-kl
        o[8] = (((4.0 * pi) * pj) * pk) * pl;
        break;
    }
}

// m_s[gx][gy] (cls.rs:6-26) over IIS_A (conditional.rs:10-50): genotype
// indices in VCF order, alleles (i, j) of gx and (k, l) of gy, their IBS mode.
__device__ __forceinline__ void m_row_generic(const double* p, int gi_x, int gj_x, int gi_y, int gj_y, double* o) {
    cond_prob(p[gi_x], p[gj_x], p[gi_y], p[gj_y], iis_mode(gi_x, gj_x, gi_y, gj_y), o);
}

// Alleles of VCF genotype index g (j(j+1)/2 + i, i <= j), g < 10.
__device__ __forceinline__ void index_alleles(int g, int* i, int* j) {
    const int jj = g >= 6 ? 3 : g >= 3 ? 2 : g >= 1 ? 1 : 0;
    *j = jj;
    *i = g - jj * (jj + 1) / 2;
}

// The biallelic m row (IIS_2) from (p, q).
__device__ __forceinline__ void m_row(double p, double q, int gx, int gy, double* o) {
    const double f[2] = {p, q};
    int i, j, k, l;
    index_alleles(gx, &i, &j);
    index_alleles(gy, &k, &l);
    m_row_generic(f, i, j, k, l, o);
}

// Row r of the reference's concatenated m rows (coefficients.rs:280-320):
// 9 per 2-allele site (frequencies af2), then the precomputed 36-row tables of
// the 3-allele sites and 100-row tables of the 4-allele sites.
struct GramRows {
    const double* af2;
    const double* m3;
    const double* m4;
    uint64_t rows2, rows3;
    __device__ __forceinline__ void row(uint64_t r, double* m) const {
        if (r < rows2) {
            const uint64_t s = r / 9;
            const int combo = (int)(r % 9);
            m_row(af2[2 * s], af2[2 * s + 1], combo / 3, combo % 3, m);
            return;
        }
        r -= rows2;
        const double* src = r < rows3 ? m3 + r * 9 : m4 + (r - rows3) * 9;
#pragma unroll
        for (int k = 0; k < 9; k++) m[k] = src[k];
    }
};

}  // namespace

// vcf.rs:217-221: a site with a missing call in any sample is dropped on the
// GT path; flag[s] = 1 when every sample is called.
__global__ __launch_bounds__(64) void complete_sites_kernel(const uint8_t* __restrict__ packed, size_t row_bytes,
                                                             uint32_t n_samples, uint32_t n_sites,
                                                             uint8_t* __restrict__ flag) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sites) return;
    uint8_t ok = 1;
    for (uint32_t i = 0; i < n_samples; i++)
        if (((packed[(size_t)i * row_bytes + (s >> 2)] >> ((s & 3u) * 2u)) & 3u) == 3u) ok = 0;
    flag[s] = ok;
}

// wrapper.rs:412-428: allele frequencies by counting over all samples.
__global__ __launch_bounds__(64) void gt_count_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                       uint32_t n_samples, uint32_t n_sites,
                                                       double* __restrict__ af) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sites) return;
    uint64_t f0 = 0, f1 = 0;
    for (uint32_t i = 0; i < n_samples; i++) {
        const uint32_t g = (geno[(size_t)i * row_words + s / 16] >> (2 * (s % 16))) & 3u;
        f0 += g == 0 ? 2 : g == 1 ? 1 : 0;
        f1 += g == 2 ? 2 : g == 1 ? 1 : 0;
    }
    const double tot = (double)(2 * (uint64_t)n_samples);
    af[2 * s] = (double)f0 / tot;
    af[2 * s + 1] = (double)f1 / tot;
}

// quadratic.rs:68-365 per tile: thread (tile t, lane j) sums its 32 rows
// r = 256 t + 8 b + j of the concatenated m rows (GramRows) from zero into
// z[t][45][8].
__global__ __launch_bounds__(64) void gram_tile_kernel(GramRows gr, uint64_t n_blocks, uint64_t n_tiles,
                                                        double* __restrict__ z) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_tiles * 8) return;
    const uint64_t t = tid / 8;
    const int j = (int)(tid % 8);
    double acc[45];
#pragma unroll
    for (int k = 0; k < 45; k++) acc[k] = 0.0;
    const uint64_t b0 = t * 32, b1 = b0 + 32 < n_blocks ? b0 + 32 : n_blocks;
    for (uint64_t b = b0; b < b1; b++) {
        double m[9];
        gr.row(b * 8 + j, m);
        int id = 0;
#pragma unroll
        for (int a = 0; a < 9; a++)
#pragma unroll
            for (int c = a; c < 9; c++, id++) acc[id] = fma(m[a], m[c], acc[id]);
    }
#pragma unroll
    for (int k = 0; k < 45; k++) z[(t * 45 + k) * 8 + j] = acc[k];
}

// Running lane sums over tiles (h = h + z_t), radd, the scalar remainder rows
// as one partial tile, divided by the number of sites (quadratic.rs:12-33).
__global__ __launch_bounds__(64) void gram_final_kernel(GramRows gr, const double* __restrict__ z, uint64_t n_tiles,
                                                         uint64_t n_blocks, uint32_t n_rem, uint32_t n_sites,
                                                         double* __restrict__ q) {
    const int id = threadIdx.x;
    if (id >= 45) return;
    int a = 0, c = 0;
    for (int k = 0, row = 0; row < 9; row++)
        for (int col = row; col < 9; col++, k++)
            if (k == id) {
                a = row;
                c = col;
            }
    double h[8];
    for (int j = 0; j < 8; j++) h[j] = 0.0;
    for (uint64_t t = 0; t < n_tiles; t++)
        for (int j = 0; j < 8; j++) h[j] = h[j] + z[(t * 45 + id) * 8 + j];
    double rz = 0.0;
    for (uint32_t r = 0; r < n_rem; r++) {
        double m[9];
        gr.row(n_blocks * 8 + r, m);
        rz = fma(m[a], m[c], rz);
    }
    const double v = (radd8(h) + (0.0 + rz)) / (double)n_sites;
    q[a * 9 + c] = v;
    q[c * 9 + a] = v;
}

// wrapper.rs:412-428 for the sites with A alleles: genotype indices in
// 4-bit rows (8 sites per word, sample-major); af n_sites x A.
template <int A>
__global__ __launch_bounds__(64) void gt_count_multi_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                             uint32_t n_samples, uint32_t n_sites,
                                                             double* __restrict__ af) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sites) return;
    uint64_t f[A];
#pragma unroll
    for (int a = 0; a < A; a++) f[a] = 0;
    for (uint32_t n = 0; n < n_samples; n++) {
        const int g = (int)((geno[(size_t)n * row_words + s / 8] >> (4 * (s % 8))) & 15u);
        int i, j;
        index_alleles(g, &i, &j);
#pragma unroll
        for (int a = 0; a < A; a++) f[a] += (uint64_t)(a == i) + (uint64_t)(a == j);
    }
    const double tot = (double)(2 * (uint64_t)n_samples);
#pragma unroll
    for (int a = 0; a < A; a++) af[(size_t)s * A + a] = (double)f[a] / tot;
}

// cls.rs:6-26 m-matrices of the sites with A alleles: m[s][gx][gy][9], one
// thread per (site, genotype pair).
template <int A>
__global__ __launch_bounds__(64) void cls_m_kernel(const double* __restrict__ af, uint32_t n_sites,
                                                    double* __restrict__ m) {
    constexpr int G = A * (A + 1) / 2;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n_sites * G * G) return;
    const uint64_t s = e / (G * G);
    const int combo = (int)(e % (G * G));
    int i, j, k, l;
    index_alleles(combo / G, &i, &j);
    index_alleles(combo % G, &k, &l);
    double o[9];
    m_row_generic(af + s * A, i, j, k, l, o);
#pragma unroll
    for (int t = 0; t < 9; t++) m[e * 9 + t] = o[t];
}

// cls.rs:28-49 over the sites with A alleles, one pair per lane: c_A = sum of
// m_s[gx][gy] in site order from zero.  The block stages S sites' m tables in
// LDS (row gx G + gy; row G G is zero and takes the padding nibbles 15).
// A = 3 stores c3; A = 4 stores c3 + c4 (cls.rs via coefficients.rs:226-228)
// with c3 read back from out, or 0 + c4 when there are no 3-allele sites.
template <int A>
__global__ __launch_bounds__(256) void cls_multi_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                         const double* __restrict__ mtab, uint32_t n_sites,
                                                         const uint32_t* __restrict__ pairs, uint64_t npairs,
                                                         int add_prev, double* __restrict__ out) {
    constexpr int G = A * (A + 1) / 2, GG = G * G, S = A == 4 ? 4 : 8, W = 10, RS = S * W + 2;
    __shared__ __attribute__((aligned(16))) double lut[(GG + 1) * RS];
    const uint32_t tid = threadIdx.x;
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + tid;
    const bool live = idx < npairs;
    for (uint32_t i = tid; i < RS; i += 256) lut[GG * RS + i] = 0.0;
    uint32_t x = 0, y = 0;
    if (live) {
        x = pairs[2 * idx];
        y = pairs[2 * idx + 1];
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    for (uint32_t cs = 0; cs < n_sites; cs += S) {
        __syncthreads();
        for (uint32_t e = tid; e < (uint32_t)(GG * S); e += 256) {
            const uint32_t k = e / GG, row = e % GG, site = cs + k;
            double* dst = lut + row * RS + k * W;
            if (site < n_sites) {
                const double* src = mtab + ((size_t)site * GG + row) * 9;
                for (int t = 0; t < 9; t++) dst[t] = src[t];
            } else {
                for (int t = 0; t < 9; t++) dst[t] = 0.0;
            }
        }
        __syncthreads();
        const uint32_t wx = rowx[cs / 8] >> (4 * (cs % 8));
        const uint32_t wy = rowy[cs / 8] >> (4 * (cs % 8));
#pragma unroll
        for (int k = 0; k < S; k++) {
            const uint32_t gx = (wx >> (4 * k)) & 15u, gy = (wy >> (4 * k)) & 15u;
            const uint32_t row = gx < (uint32_t)G ? gx * G + gy : (uint32_t)GG;
            const double* r = lut + row * RS + k * W;
#pragma unroll
            for (int t = 0; t < 9; t++) c[t] = c[t] + r[t];
        }
    }
    if (!live) return;
    double* o = out + idx * 9;
#pragma unroll
    for (int t = 0; t < 9; t++) o[t] = A == 4 ? (add_prev ? o[t] : 0.0) + c[t] : c[t];
}

// cls.rs:28-49 over the 2-allele sites, one pair per lane: c2 = sum of
// m_s[gx][gy] in site order from zero.  The block stages a 32-site chunk's
// 12-row table in LDS (row min(4 gx + gy, 11); rows 3, 7, 11 and padding sites
// are zero: x + 0 = x) and each lane reads its 72-byte row per site.
__global__ __launch_bounds__(256) void cls_sum_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                       const double* __restrict__ af, uint32_t n_sites,
                                                       const uint32_t* __restrict__ pairs, uint64_t npairs,
                                                       double* __restrict__ csum) {
    __shared__ __attribute__((aligned(16))) double lut[LUT_BUF];
    const uint32_t tid = threadIdx.x;
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + tid;
    const bool live = idx < npairs;
    for (uint32_t i = tid; i < LUT_BUF; i += 256) lut[i] = 0.0;
    uint32_t x = 0, y = 0;
    if (live) {
        x = pairs[2 * idx];
        y = pairs[2 * idx + 1];
    }
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    // (site, row) of the table rows this thread builds: the 32 sites of a chunk
    // across the lanes and the rows across the wavefronts, so that a
    // wavefront's lanes take two of lut_block's cases (three for wavefront 0,
    // which adds row (2,2)) instead of all nine
    const int lut_site = tid & (CHUNK - 1), lut_part = tid / CHUNK;
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    const uint32_t v_pad = row_words * 16;
    for (uint32_t cs = 0; cs < v_pad; cs += CHUNK) {
        __syncthreads();  // the previous chunk's rows are read
        {
            const uint32_t s = cs + lut_site;
            const int parts[2] = {lut_part, 8};
            for (int w = 0; w < (lut_part == 0 ? 2 : 1); w++) {
                const int pr = parts[w];
                const int gx = pr / 3, gy = pr % 3;
                double* dst = lut + (gx * 4 + gy) * LUT_RS + lut_site * LUT_W;
                if (s < n_sites) m_row(af[2 * s], af[2 * s + 1], gx, gy, dst);
                else
                    for (int k = 0; k < 9; k++) dst[k] = 0.0;
            }
        }
        const uint2 wx = *reinterpret_cast<const uint2*>(rowx + cs / 16);
        const uint2 wy = *reinterpret_cast<const uint2*>(rowy + cs / 16);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            const uint32_t ax = (k < 16 ? wx.x : wx.y) >> (2 * (k & 15));
            const uint32_t ay = (k < 16 ? wy.x : wy.y) >> (2 * (k & 15));
            const uint32_t row = min(((ax & 3u) << 2) | (ay & 3u), 11u);
            const double* r = lut + row * LUT_RS + k * LUT_W;
            const double2* r2 = reinterpret_cast<const double2*>(r);
            double m[9];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const double2 v = r2[j];
                m[2 * j] = v.x;
                m[2 * j + 1] = v.y;
            }
            m[8] = r[8];
#pragma unroll
            for (int j = 0; j < 9; j++) c[j] = c[j] + m[j];
        }
    }
    if (!live) return;
#pragma unroll
    for (int k = 0; k < 9; k++) csum[idx * 9 + k] = c[k];
}

// coefficients.rs:216-276 for one pair per lane: c = (c2 + (c3 + c4)) / -V
// (coefficients.rs:226-228) from the per-bucket sums, the QP from 1/9 (a self
// pair solves the 2-variable (delta1, delta7) QP from (1/2, 1/2)), brier,
// gradient and condition.
__global__ __launch_bounds__(64) void cls_pairs_kernel(const double* __restrict__ csum,
                                                        const double* __restrict__ c34, uint32_t v_total,
                                                        const double* __restrict__ qmat,
                                                        const uint32_t* __restrict__ pairs, uint64_t npairs,
                                                        Tune tune, kestrel_output* __restrict__ out) {
    const uint64_t idx = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    if (idx >= npairs) return;
    const uint32_t x = pairs[2 * idx], y = pairs[2 * idx + 1];
    const double nv = -(double)v_total;
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = (csum[idx * 9 + k] + (c34 ? c34[idx * 9 + k] : 0.0 + 0.0)) / nv;
    double Q[9][9];
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int j = 0; j < 9; j++) Q[i][j] = qmat[i * 9 + j];
    double delta[9];
    uint64_t iters;
    if (x == y) {
        const double Q2[2][2] = {{Q[0][0], Q[0][6]}, {Q[6][0], Q[6][6]}};
        const double c2[2] = {c[0], c[6]};
        const double y0[2] = {0.5, 0.5};
        double y2[2];
        iters = solve_qp_reg<2>(Q2, c2, y0, tune, y2);
#pragma unroll
        for (int k = 0; k < 9; k++) delta[k] = 0.0;
        delta[0] = y2[0];
        delta[6] = y2[1];
    } else {
        double y0[9];
#pragma unroll
        for (int k = 0; k < 9; k++) y0[k] = 1.0 / 9.0;
        iters = solve_qp_reg<9>(Q, c, y0, tune, delta);
    }
    double qd[9], g[9], dq = 0.0, dc = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 9; j++) s = fma(Q[i][j], delta[j], s);
        qd[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) dq = fma(delta[i], qd[i], dq);
#pragma unroll
    for (int i = 0; i < 9; i++) dc = fma(c[i], delta[i], dc);
    const double brier = dq + 2.0 * dc + 1.0;
#pragma unroll
    for (int i = 0; i < 9; i++) g[i] = qd[i] + c[i];
    kestrel_output o;
    o.x = x;
    o.y = y;
#pragma unroll
    for (int k = 0; k < 9; k++) o.jacquard[k] = delta[k];
    o.obj = brier;
    o.iters = iters;
    o.cond = condition_reg<9>(delta, g, Q, tune.sqp_conv_tol);
    o.n_sites = v_total;
    o.status = KESTREL_PAIR_OK;
    out[idx] = o;
}

// ------------------------------------------------------------ launchers

namespace klaunch {

hipError_t af(const uint8_t* packed, size_t row_bytes, uint32_t n_samples, uint32_t n_sites, const Tune& t,
              double maf, double* af_out, int* status, hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    hipLaunchKernelGGL(af_kernel, dim3((n_sites + 63) / 64), dim3(64), 0, st, packed, row_bytes, n_samples, n_sites,
                       t, maf, af_out, status);
    return hipGetLastError();
}

hipError_t gather(const uint8_t* src, size_t src_row_bytes, const uint32_t* kept, uint32_t n_kept,
                  uint32_t n_samples, uint32_t* dst, uint32_t dst_row_words, hipStream_t st, const uint32_t* dst_rows) {
    const size_t n = (size_t)n_samples * dst_row_words;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, src_row_bytes, kept,
                       n_kept, n_samples, dst, dst_row_words, dst_rows);
    return hipGetLastError();
}

hipError_t sweep(const uint32_t* geno, uint32_t row_words, const double* af, const uint32_t* seg_bounds,
                 uint32_t nseg, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active,
                 uint32_t nslots, double eps, double* partial, hipStream_t st, const double* gl, uint32_t v_pad,
                 int gl_stride, bool skip_masked, bool mixed) {
    dim3 grid((nslots + SWEEP_TPB - 1) / SWEEP_TPB, nseg);
    if (gl && mixed && gl_stride == 2)
        hipLaunchKernelGGL(sweep_gl_mixed<2>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (gl && mixed)
        hipLaunchKernelGGL(sweep_gl_mixed<3>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (gl && gl_stride == 2)
        hipLaunchKernelGGL(sweep_gl<2>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af, seg_bounds, sx,
                           sy, pt, active, nslots, eps, partial);
    else if (gl)
        hipLaunchKernelGGL(sweep_gl<3>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af, seg_bounds, sx,
                           sy, pt, active, nslots, eps, partial);
    else if (mixed && skip_masked)
        hipLaunchKernelGGL(sweep_mixed<true>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, af, seg_bounds, sx, sy,
                           pt, active, nslots, eps, partial);
    else if (mixed)
        hipLaunchKernelGGL(sweep_mixed<false>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, af, seg_bounds, sx, sy,
                           pt, active, nslots, eps, partial);
    else if (skip_masked)
        hipLaunchKernelGGL(sweep_fp64<true>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, af, seg_bounds, sx, sy,
                           pt, active, nslots, eps, partial);
    else
        hipLaunchKernelGGL(sweep_fp64<false>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, af, seg_bounds, sx, sy,
                           pt, active, nslots, eps, partial);
    return hipGetLastError();
}

hipError_t sweep_lanes(const uint32_t* geno, uint32_t row_words, const double* af, const uint32_t* seg_bounds,
                       uint32_t nseg, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active,
                       uint32_t nslots, double eps, double* partial, hipStream_t st, const double* gl,
                       uint32_t v_pad, int gl_stride) {
    constexpr uint32_t per_block = SWEEP_TPB / 64;
    dim3 grid((nslots + per_block - 1) / per_block, nseg);
    if (gl && gl_stride == 2)
        hipLaunchKernelGGL(sweep_lanes_gl<2>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (gl)
        hipLaunchKernelGGL(sweep_lanes_gl<3>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else
        hipLaunchKernelGGL(sweep_lanes_fp64, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, af, seg_bounds, sx, sy,
                           pt, active, nslots, eps, partial);
    return hipGetLastError();
}

hipError_t af_gl(const double* gl, uint32_t n_samples, uint32_t n_sites, const Tune& t, double maf, double* af_out,
                 int* status, hipStream_t st, int alleles) {
    if (n_sites == 0) return hipSuccess;
    const dim3 grid((n_sites + 63) / 64);
    if (alleles == 2)
        hipLaunchKernelGGL(af_gl_kernel<2>, grid, dim3(64), 0, st, gl, n_samples, n_sites, t, maf, af_out, status);
    else if (alleles == 3)
        hipLaunchKernelGGL(af_gl_kernel<3>, grid, dim3(64), 0, st, gl, n_samples, n_sites, t, maf, af_out, status);
    else if (alleles == 4)
        hipLaunchKernelGGL(af_gl_kernel<4>, grid, dim3(64), 0, st, gl, n_samples, n_sites, t, maf, af_out, status);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t sweep_multi(int alleles, const uint32_t* geno, uint32_t row_words, const double* gl, uint32_t v_pad,
                       const double* table, const double* af, const uint32_t* seg_bounds, uint32_t nseg,
                       const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active, uint32_t nslots,
                       double eps, double* partial, hipStream_t st, const float* table32, const uint32_t* codes,
                       const double* lut) {
    dim3 grid((nslots + SWEEP_TPB - 1) / SWEEP_TPB, nseg);
    // hard calls: the genotype pairs' rows (the same bits); blocks of 256
    // pairs when blocks of 512 would not give every CU one
    int dev = 0, cus = 0;
    if (codes && !table32) {
        const hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        const hipError_t e2 = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e2 != hipSuccess) return e2;
    }
    const bool half = (uint64_t)((nslots + GLC_TPB - 1) / GLC_TPB) * nseg < (uint64_t)cus;
    const dim3 gcodes((nslots + GLC_TPB - 1) / GLC_TPB, nseg), ghalf((nslots + 255) / 256, nseg);
    if (codes && !table32 && alleles == 3 && !half)
        hipLaunchKernelGGL((sweep_glm_codes<3, GLC_TPB>), gcodes, dim3(GLC_TPB), 0, st, geno, row_words, codes, lut,
                           seg_bounds, sx, sy, pt, active, nslots, v_pad, eps, partial);
    else if (codes && !table32 && alleles == 3)
        hipLaunchKernelGGL((sweep_glm_codes<3, 256>), ghalf, dim3(256), 0, st, geno, row_words, codes, lut,
                           seg_bounds, sx, sy, pt, active, nslots, v_pad, eps, partial);
    else if (codes && !table32 && alleles == 4 && !half)
        hipLaunchKernelGGL((sweep_glm_codes<4, GLC_TPB>), gcodes, dim3(GLC_TPB), 0, st, geno, row_words, codes, lut,
                           seg_bounds, sx, sy, pt, active, nslots, v_pad, eps, partial);
    else if (codes && !table32 && alleles == 4)
        hipLaunchKernelGGL((sweep_glm_codes<4, 256>), ghalf, dim3(256), 0, st, geno, row_words, codes, lut,
                           seg_bounds, sx, sy, pt, active, nslots, v_pad, eps, partial);
    else if (table32 && alleles == 3)
        hipLaunchKernelGGL(sweep_glm_mixed<3>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, table32,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (table32 && alleles == 4)
        hipLaunchKernelGGL(sweep_glm_mixed<4>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, table32,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (table32)
        return hipErrorInvalidValue;
    else if (alleles == 3)  // coefficients in registers: 1.1-1.2x the table form (LDS-bound), the same bits
        hipLaunchKernelGGL((sweep_glm<3, true>), grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, af,
                           seg_bounds, sx, sy, pt, active, nslots, eps, partial);
    else if (alleles == 4)  // in registers the products shared between pairs spill (46 VGPRs)
        hipLaunchKernelGGL(sweep_glm<4>, grid, dim3(SWEEP_TPB), 0, st, geno, row_words, gl, v_pad, table, seg_bounds,
                           sx, sy, pt, active, nslots, eps, partial);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t sweep_exact(int kind, const uint32_t* geno, uint32_t row_words, const double* gl, uint32_t v_pad,
                       const double* af_or_table, const uint32_t* seg_bounds, uint32_t nseg, const uint32_t* sx,
                       const uint32_t* sy, const double* pt, const int* exact, uint32_t nslots, double eps,
                       double* partial, hipStream_t st) {
    const dim3 grid((nslots + 63) / 64, nseg), block(64);
    switch (kind) {
    case 0:
        hipLaunchKernelGGL(sweep_exact_kernel<0>, grid, block, 0, st, geno, row_words, gl, v_pad, af_or_table, seg_bounds,
                           sx, sy, pt, exact, nslots, eps, partial);
        break;
    case 2:
        hipLaunchKernelGGL(sweep_exact_kernel<2>, grid, block, 0, st, geno, row_words, gl, v_pad, af_or_table, seg_bounds,
                           sx, sy, pt, exact, nslots, eps, partial);
        break;
    case 3:
        hipLaunchKernelGGL(sweep_exact_kernel<3>, grid, block, 0, st, geno, row_words, gl, v_pad, af_or_table, seg_bounds,
                           sx, sy, pt, exact, nslots, eps, partial);
        break;
    case 13:
        hipLaunchKernelGGL(sweep_exact_kernel<13>, grid, block, 0, st, geno, row_words, gl, v_pad, af_or_table, seg_bounds,
                           sx, sy, pt, exact, nslots, eps, partial);
        break;
    case 14:
        hipLaunchKernelGGL(sweep_exact_kernel<14>, grid, block, 0, st, geno, row_words, gl, v_pad, af_or_table, seg_bounds,
                           sx, sy, pt, exact, nslots, eps, partial);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t part_codes(int alleles, const double* rows, const uint32_t* mask, uint32_t row_words, uint32_t v_pad,
                      uint32_t n_samples, uint32_t* codes, int* bad, hipStream_t st) {
    const uint64_t n = (uint64_t)n_samples * (v_pad / 8);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (alleles == 3)
        hipLaunchKernelGGL(part_codes_kernel<3>, grid, dim3(256), 0, st, rows, mask, row_words, v_pad, n_samples, codes,
                           bad);
    else if (alleles == 4)
        hipLaunchKernelGGL(part_codes_kernel<4>, grid, dim3(256), 0, st, rows, mask, row_words, v_pad, n_samples, codes,
                           bad);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

size_t part_lut_size(int alleles) {
    const size_t G = (size_t)alleles * (alleles + 1) / 2;
    return (G * G + 1) * GLC_RW;
}

hipError_t part_lut(int alleles, const double* table, uint32_t v_pad, double* lut, hipStream_t st) {
    const uint64_t n = (uint64_t)v_pad * (part_lut_size(alleles) / GLC_RW);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (alleles == 3)
        hipLaunchKernelGGL(part_lut_kernel<3>, grid, dim3(256), 0, st, table, v_pad, lut);
    else if (alleles == 4)
        hipLaunchKernelGGL(part_lut_kernel<4>, grid, dim3(256), 0, st, table, v_pad, lut);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t any_nonzero(const double* rows, uint64_t n, int* nonzero, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(any_nonzero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rows, n, nonzero);
    return hipGetLastError();
}

hipError_t round_f32(const double* src, float* dst, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(round_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

int iis_table_size(int alleles) {
    return alleles == 3 ? IisTable<3>().off[36] : alleles == 4 ? IisTable<4>().off[100] : 0;
}

hipError_t iis_table(int alleles, const double* af, uint32_t n_sites, double* table, hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    const uint64_t gg = alleles == 3 ? 36 : 100, n = (uint64_t)n_sites * gg;
    const dim3 grid((unsigned)((n + 63) / 64));
    if (alleles == 3)
        hipLaunchKernelGGL(iis_table_kernel<3>, grid, dim3(64), 0, st, af, n_sites, table);
    else if (alleles == 4)
        hipLaunchKernelGGL(iis_table_kernel<4>, grid, dim3(64), 0, st, af, n_sites, table);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t gather_multi(const double* src2, const double* src3, const double* src4, uint32_t n_samples,
                        const uint32_t* map_site, const uint8_t* map_alleles, const uint8_t* gmap, uint32_t n_t,
                        int gp, double* rows, uint32_t v_pad, uint32_t* mask, uint32_t row_words, hipStream_t st,
                        int compact, int* no_one) {
    const size_t n = (size_t)n_samples * row_words;
    if (n == 0) return hipSuccess;
    if (compact && gp != 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_multi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src2, src3, src4,
                       n_samples, map_site, map_alleles, gmap, n_t, gp, rows, v_pad, mask, row_words, compact, no_one);
    return hipGetLastError();
}

hipError_t count_missing(const uint32_t* geno, uint32_t n_samples, uint32_t row_words, uint32_t n_sites,
                         unsigned long long* count, hipStream_t st) {
    const size_t n = (size_t)n_samples * row_words;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(count_missing_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, geno, n_samples,
                       row_words, n_sites, count);
    return hipGetLastError();
}

hipError_t gather_gl(const double* gl_site, uint32_t n_samples, const uint32_t* kept, uint32_t n_kept, double* rows,
                     uint32_t v_pad, uint32_t* mask, uint32_t row_words, int compact, int* no_one, hipStream_t st) {
    const size_t n = (size_t)n_samples * row_words;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_gl_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gl_site, n_samples, kept,
                       n_kept, rows, v_pad, mask, row_words, compact, no_one);
    return hipGetLastError();
}

hipError_t step(SlotState* stt, uint32_t* sx, uint32_t* sy, double* pt, int* active, uint32_t nslots,
                double* partial, uint32_t nseg, const uint32_t* pairs, uint64_t npairs, const uint64_t* order,
                unsigned long long* next_pair, kestrel_output* out, const Tune& t, int* n_live, const PairQueue& q,
                int* live_out, hipStream_t st, const RunClaim& rc) {
    if (nseg > 1) {
        const size_t n = (size_t)NPART * nslots;
        hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, partial, nseg, nslots,
                           active);
    }
    if (rc.runs && !q.cap)
        hipLaunchKernelGGL(step_runs_kernel, dim3((nslots + SWEEP_TPB - 1) / SWEEP_TPB), dim3(SWEEP_TPB), 0, st, stt,
                           sx, sy, pt, active, nslots, partial, nseg, pairs, npairs, order, next_pair, out, t, n_live,
                           live_out, rc);
    else
        hipLaunchKernelGGL(step_kernel, dim3((nslots + 63) / 64), dim3(64), 0, st, stt, sx, sy, pt, active, nslots,
                           partial, nseg, pairs, npairs, order, next_pair, out, t, n_live, q, live_out);
    return hipGetLastError();
}

hipError_t selftest(uint64_t n, uint64_t seed, unsigned long long* bad, hipStream_t st) {
    hipLaunchKernelGGL(selftest_kernel, dim3(4096), dim3(256), 0, st, n, seed, bad);
    return hipGetLastError();
}

hipError_t calib_lds(uint32_t blocks, uint32_t iters, double* sink, hipStream_t st) {
    hipLaunchKernelGGL(calib_lds_kernel, dim3(blocks), dim3(256), 0, st, iters, sink);
    return hipGetLastError();
}

hipError_t calib_fp32(uint32_t blocks, uint32_t iters, float* sink, hipStream_t st) {
    hipLaunchKernelGGL(calib_fma_kernel<float>, dim3(blocks), dim3(256), 0, st, iters, 0.999999f, 1e-7f, sink);
    return hipGetLastError();
}

hipError_t calib_fp64(uint32_t blocks, uint32_t iters, double* sink, hipStream_t st) {
    hipLaunchKernelGGL(calib_fma_kernel<double>, dim3(blocks), dim3(256), 0, st, iters, 0.999999, 1e-7, sink);
    return hipGetLastError();
}

hipError_t copy(const void* src, void* dst, uint64_t n16, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

hipError_t complete_sites(const uint8_t* packed, size_t row_bytes, uint32_t n_samples, uint32_t n_sites, uint8_t* flag,
                          hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    hipLaunchKernelGGL(complete_sites_kernel, dim3((n_sites + 63) / 64), dim3(64), 0, st, packed, row_bytes, n_samples,
                       n_sites, flag);
    return hipGetLastError();
}

hipError_t gt_count(const uint32_t* geno, uint32_t row_words, uint32_t n_samples, uint32_t n_sites, double* af,
                    hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    hipLaunchKernelGGL(gt_count_kernel, dim3((n_sites + 63) / 64), dim3(64), 0, st, geno, row_words, n_samples,
                       n_sites, af);
    return hipGetLastError();
}

hipError_t gram(const double* af2, uint32_t n2, const double* m3, uint32_t n3, const double* m4, uint32_t n4,
                double* z, double* q, hipStream_t st) {
    const GramRows gr{af2, m3, m4, (uint64_t)n2 * 9, (uint64_t)n3 * 36};
    const uint64_t rows = gr.rows2 + gr.rows3 + (uint64_t)n4 * 100, n_blocks = rows / 8, n_tiles = (n_blocks + 31) / 32;
    const uint32_t n_rem = (uint32_t)(rows % 8);
    if (n_tiles)
        hipLaunchKernelGGL(gram_tile_kernel, dim3((unsigned)((n_tiles * 8 + 63) / 64)), dim3(64), 0, st, gr, n_blocks,
                           n_tiles, z);
    hipLaunchKernelGGL(gram_final_kernel, dim3(1), dim3(64), 0, st, gr, z, n_tiles, n_blocks, n_rem, n2 + n3 + n4, q);
    return hipGetLastError();
}

uint64_t gram_tiles(uint32_t n2, uint32_t n3, uint32_t n4) {
    const uint64_t rows = (uint64_t)n2 * 9 + (uint64_t)n3 * 36 + (uint64_t)n4 * 100;
    return (rows / 8 + 31) / 32;
}

hipError_t gt_count_multi(int alleles, const uint32_t* geno, uint32_t row_words, uint32_t n_samples, uint32_t n_sites,
                          double* af, hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    const dim3 grid((n_sites + 63) / 64);
    if (alleles == 3)
        hipLaunchKernelGGL(gt_count_multi_kernel<3>, grid, dim3(64), 0, st, geno, row_words, n_samples, n_sites, af);
    else
        hipLaunchKernelGGL(gt_count_multi_kernel<4>, grid, dim3(64), 0, st, geno, row_words, n_samples, n_sites, af);
    return hipGetLastError();
}

hipError_t cls_m(int alleles, const double* af, uint32_t n_sites, double* m, hipStream_t st) {
    if (n_sites == 0) return hipSuccess;
    const uint64_t gg = alleles == 3 ? 36 : 100, n = (uint64_t)n_sites * gg;
    const dim3 grid((unsigned)((n + 63) / 64));
    if (alleles == 3)
        hipLaunchKernelGGL(cls_m_kernel<3>, grid, dim3(64), 0, st, af, n_sites, m);
    else
        hipLaunchKernelGGL(cls_m_kernel<4>, grid, dim3(64), 0, st, af, n_sites, m);
    return hipGetLastError();
}

hipError_t cls_multi(int alleles, const uint32_t* geno, uint32_t row_words, const double* mtab, uint32_t n_sites,
                     const uint32_t* pairs, uint64_t npairs, int add_prev, double* out, hipStream_t st) {
    if (npairs == 0) return hipSuccess;
    const dim3 grid((unsigned)((npairs + 255) / 256));
    if (alleles == 3)
        hipLaunchKernelGGL(cls_multi_kernel<3>, grid, dim3(256), 0, st, geno, row_words, mtab, n_sites, pairs, npairs,
                           add_prev, out);
    else
        hipLaunchKernelGGL(cls_multi_kernel<4>, grid, dim3(256), 0, st, geno, row_words, mtab, n_sites, pairs, npairs,
                           add_prev, out);
    return hipGetLastError();
}

hipError_t cls_pairs(const uint32_t* geno, uint32_t row_words, const double* af, uint32_t n_sites, const double* c34,
                     uint32_t v_total, const double* q, const uint32_t* pairs, uint64_t npairs, const Tune& t,
                     double* csum, kestrel_output* out, hipStream_t st) {
    if (npairs == 0) return hipSuccess;
    hipLaunchKernelGGL(cls_sum_kernel, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, st, geno, row_words, af,
                       n_sites, pairs, npairs, csum);
    hipLaunchKernelGGL(cls_pairs_kernel, dim3((unsigned)((npairs + 63) / 64)), dim3(64), 0, st, csum, c34, v_total, q,
                       pairs, npairs, t, out);
    return hipGetLastError();
}

hipError_t finalize(const double* partial, uint32_t nseg, uint32_t nslots, uint32_t npts, double* f, double* g,
                    double* h, uint32_t* nsites, hipStream_t st) {
    hipLaunchKernelGGL(finalize_eval, dim3((npts + 63) / 64), dim3(64), 0, st, partial, nseg, nslots, npts, f, g, h,
                       nsites);
    return hipGetLastError();
}

hipError_t recondition(const double* partial, uint32_t nseg, uint32_t nslots, uint32_t npts, const double* pt,
                       double tol, double* cond, hipStream_t st) {
    hipLaunchKernelGGL(recondition_kernel, dim3((npts + 63) / 64), dim3(64), 0, st, partial, nseg, nslots, npts, pt,
                       tol, cond);
    return hipGetLastError();
}

}  // namespace klaunch
