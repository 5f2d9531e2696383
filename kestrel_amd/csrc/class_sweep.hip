// class_sweep.hip -- the class-sorted FP64 sweep of 2-bit cohorts: one pair
// per wavefront, its usable sites sorted by genotype-pair pattern, the sums
// carried in each pattern's small basis.
//
// c_s = mixture_row2(one-hot gx, one-hot gy; p, q) (mixture.rs:96-236) has at
// most four distinct non-zero values, in positions fixed by (gx, gy)
// (mixture.rs:137-227):
//   HS (0,0),(2,2): c = (a, a^2, a^2, a^3, a^2, a^3, a^2, a^3, a^4), a = p | q
//   B  (0,1),(2,1): c_{2,3,7,8} = u (1, 2w, w, 2w^2),   u = pq, w = p | q
//   C  (1,0),(1,2): c_{4,5,7,8} = u (1, 2w, w, 2w^2)
//   D  (0,2),(2,0): c_{1,3,5,8} = u (1, w1, w2, w1 w2), (w1, w2) = (q, p) | (p, q)
//   E  (1,1):       c_{6,7,8}   = (2u, u, 4u^2)
// so over the sites of one pattern L = eps + c.D is a Horner polynomial of
// 2-4 fma, and the gradient (grad_hess.rs:122-126) and Hessian (149-448)
// terms are 2-4 and 3-10 sums instead of 9 and 45.  The 47 sums of a pair are
// expanded into f, g[9], h[45] at the end of its segment, in the layout of
// sweep_fp64's partials (the fold and the step are unchanged).
//
// class_build_kernel sorts a pair's sites once, when its slot takes it: per
// block of CLASS_BLOCK sites, the list of each pattern's sites in ascending
// order (entries (site in block << 1) | flag, flag = the pattern's second
// genotype pair) and the five list lengths.  sweep_class_fp64 then stages the
// block's frequencies in LDS for the workgroup's pairs and walks the lists:
// entry j of a pattern's list is lane j % 64's site in chunk j / 64.
//
// The summation order (each lane's sums in list order, a log of the lane's
// product after every 16th chunk of the segment and at its end, a butterfly
// over the 64 lanes, the expansion) is restated by the oracle's class order;
// records are bit-identical to it, and within rounding of the reference's
// order (different from sweep_fp64's in the last bits).
#include "kestrel_internal.hpp"

using namespace kdev;

namespace {

constexpr int NACC = 47;  // objective + the pattern sums

// nibble per genotype pair (4 gx + gy): pattern (0 HS, 1 B, 2 C, 3 D, 4 E,
// 7 missing) | flag << 3
constexpr uint64_t PATTERN_LUT =
    (0x0ull << 0) | (0x1ull << 4) | (0x3ull << 8) | (0x7ull << 12) |   // (0,0) (0,1) (0,2) (0,3)
    (0x2ull << 16) | (0x4ull << 20) | (0xAull << 24) | (0x7ull << 28) |  // (1,0) (1,1) (1,2) (1,3)
    (0xBull << 32) | (0x9ull << 36) | (0x8ull << 40) | (0x7ull << 44) |  // (2,0) (2,1) (2,2) (2,3)
    (0x7ull << 48) | (0x7ull << 52) | (0x7ull << 56) | (0x7ull << 60);   // (3,*)

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// inclusive sum over lanes 0..lane
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

}  // namespace

// One wavefront per slot whose pair (sx, sy) differs from the one its lists
// were built for (built[slot]); four a workgroup.  Lane l takes the sites
// 64 l .. 64 l + 63 of each block (four words of each row), counts its sites
// of each pattern, a wavefront scan gives each lane its positions, and the
// lane writes its entries in site order through per-lane LDS counters.
__global__ __launch_bounds__(256) void class_build_kernel(const uint32_t* __restrict__ geno, uint32_t row_words,
                                                          const uint32_t* __restrict__ sx,
                                                          const uint32_t* __restrict__ sy,
                                                          const int* __restrict__ active, uint32_t nslots,
                                                          uint32_t nblk, uint16_t* __restrict__ lists,
                                                          uint16_t* __restrict__ counts, uint2* __restrict__ built) {
    __shared__ uint32_t ctr[4][5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t stage[4][CLASS_BLOCK];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t slot = blockIdx.x * 4 + w;
    if (slot >= nslots || active[slot] == 0) return;  // uniform over the wavefront
    const uint32_t x = sx[slot], y = sy[slot];
    const uint2 b0 = built[slot];
    if (b0.x == x && b0.y == y) return;
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    uint32_t* c = ctr[w];
    uint16_t* sg = stage[w];
    for (uint32_t b = 0; b < nblk; b++) {
        const uint32_t w0 = b * (CLASS_BLOCK / 16) + 4 * lane;
        uint4 wx = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu), wy = wx;
        if (w0 < row_words) {  // row_words is a multiple of 16: the four words are in the row
            wx = *reinterpret_cast<const uint4*>(rowx + w0);
            wy = *reinterpret_cast<const uint4*>(rowy + w0);
        }
        const uint32_t ax[4] = {wx.x, wx.y, wx.z, wx.w}, ay[4] = {wy.x, wy.y, wy.z, wy.w};
        uint32_t cnt[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t m = 0x55555555u;
            const uint32_t xl = ax[q] & m, xh = (ax[q] >> 1) & m, yl = ay[q] & m, yh = (ay[q] >> 1) & m;
            const uint32_t x0 = ~(xl | xh) & m, x1 = xl & ~xh, x2 = xh & ~xl;
            const uint32_t y0 = ~(yl | yh) & m, y1 = yl & ~yh, y2 = yh & ~yl;
            cnt[0] += __popc((x0 & y0) | (x2 & y2));
            cnt[1] += __popc(y1 & (x0 | x2));
            cnt[2] += __popc(x1 & (y0 | y2));
            cnt[3] += __popc((x0 & y2) | (x2 & y0));
            cnt[4] += __popc(x1 & y1);
        }
        // positions: pattern k's list starts after the lists of patterns < k
        const uint32_t s01 = wave_scan(cnt[0] | (cnt[1] << 16), lane);
        const uint32_t s23 = wave_scan(cnt[2] | (cnt[3] << 16), lane);
        const uint32_t s4 = wave_scan(cnt[4], lane);
        const uint32_t t01 = (uint32_t)__builtin_amdgcn_readlane((int)s01, 63);
        const uint32_t t23 = (uint32_t)__builtin_amdgcn_readlane((int)s23, 63);
        const uint32_t t4 = (uint32_t)__builtin_amdgcn_readlane((int)s4, 63);
        const uint32_t tot[5] = {t01 & 0xffffu, t01 >> 16, t23 & 0xffffu, t23 >> 16, t4};
        const uint32_t inc[5] = {s01 & 0xffffu, s01 >> 16, s23 & 0xffffu, s23 >> 16, s4};
        uint32_t base = 0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            c[k * 64 + lane] = base + inc[k] - cnt[k];
            base += tot[k];
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t idx = (((ax[q] >> (2 * i)) & 3u) << 2) | ((ay[q] >> (2 * i)) & 3u);
                const uint32_t nib = (uint32_t)(PATTERN_LUT >> (4 * idx)) & 15u;
                const uint32_t k = nib & 7u;
                if (k != 7u) {
                    const uint32_t pos = atomicAdd(&c[k * 64 + lane], 1u);
                    sg[pos] = (uint16_t)(((64u * lane + 16u * q + (uint32_t)i) << 1) | (nib >> 3));
                }
            }
        wave_sync();
        // the block's entries (base of them) to the slot's lists, 8 a lane-step
        uint16_t* dst = lists + ((size_t)slot * nblk + b) * CLASS_BLOCK;
        const uint32_t units = (base + 7) / 8;
        for (uint32_t u = lane; u < units; u += 64)
            reinterpret_cast<uint4*>(dst)[u] = reinterpret_cast<const uint4*>(sg)[u];
        if (lane == 0) {
            uint4 cv;
            cv.x = tot[0] | (tot[1] << 16);
            cv.y = tot[2] | (tot[3] << 16);
            cv.z = tot[4];
            cv.w = 0;
            reinterpret_cast<uint4*>(counts)[(size_t)slot * nblk + b] = cv;
        }
        wave_sync();
    }
    if (lane == 0) built[slot] = make_uint2(x, y);
}

namespace {

// a value equal in every lane, kept in scalar registers
__device__ __forceinline__ double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __hiloint2double(hi, lo);
}

struct ClassCoef {
    double x0, x1, x2, x3, x4, x5, x8, T1, T3, B37, B8, C57, E67, E8;
};

__device__ __forceinline__ ClassCoef class_coef(const double* D) {
    ClassCoef c;
    c.x0 = D[0];
    c.x1 = D[1];
    c.x2 = D[2];
    c.x3 = D[3];
    c.x4 = D[4];
    c.x5 = D[5];
    c.x8 = D[8];
    c.T1 = ((D[1] + D[2]) + D[4]) + D[6];
    c.T3 = (D[3] + D[5]) + D[7];
    c.B37 = (D[3] + D[3]) + D[7];
    c.B8 = D[8] + D[8];
    c.C57 = (D[5] + D[5]) + D[7];
    c.E67 = (D[6] + D[6]) + D[7];
    c.E8 = c.B8 + c.B8;
    // one pair per wavefront: the coefficients are the same in every lane
    c.x0 = uniform(c.x0);
    c.x1 = uniform(c.x1);
    c.x2 = uniform(c.x2);
    c.x3 = uniform(c.x3);
    c.x4 = uniform(c.x4);
    c.x5 = uniform(c.x5);
    c.x8 = uniform(c.x8);
    c.T1 = uniform(c.T1);
    c.T3 = uniform(c.T3);
    c.B37 = uniform(c.B37);
    c.B8 = uniform(c.B8);
    c.C57 = uniform(c.C57);
    c.E67 = uniform(c.E67);
    c.E8 = uniform(c.E8);
    return c;
}

// One site of pattern K in a lane: L into the log product, then the
// pattern's gradient and Hessian sums (A[1..46]; the layout of the oracle's
// class order).
template <int K>
__device__ __forceinline__ void class_site(double* A, double& P, uint32_t f, double p, double q, const ClassCoef& c,
                                           double eps) {
    if constexpr (K == 0) {
        const double a = f ? q : p;
        const double L = fma(a, fma(a, fma(a, fma(a, c.x8, c.T3), c.T1), c.x0), eps);
        P = P * L;
        const double r = recip_rn(L);
        const double v0 = a * r, v1 = v0 * a, v2 = v1 * a, v3 = v2 * a;
        A[1] = A[1] + v0;
        A[2] = A[2] + v1;
        A[3] = A[3] + v2;
        A[4] = A[4] + v3;
        A[5] = fma(v0, v0, A[5]);
        A[6] = fma(v0, v1, A[6]);
        A[7] = fma(v1, v1, A[7]);
        A[8] = fma(v1, v2, A[8]);
        A[9] = fma(v2, v2, A[9]);
        A[10] = fma(v2, v3, A[10]);
        A[11] = fma(v3, v3, A[11]);
    } else if constexpr (K == 1 || K == 2) {
        constexpr int o = K == 1 ? 12 : 20;
        const double u = p * q, w = f ? q : p;
        const double lin = K == 1 ? c.B37 : c.C57, con = K == 1 ? c.x2 : c.x4;
        const double L = fma(u, fma(w, fma(w, c.B8, lin), con), eps);
        P = P * L;
        const double r = recip_rn(L);
        const double b0 = u * r, b1 = b0 * w, b2 = b1 * w;
        A[o] = A[o] + b0;
        A[o + 1] = A[o + 1] + b1;
        A[o + 2] = A[o + 2] + b2;
        A[o + 3] = fma(b0, b0, A[o + 3]);
        A[o + 4] = fma(b0, b1, A[o + 4]);
        A[o + 5] = fma(b1, b1, A[o + 5]);
        A[o + 6] = fma(b1, b2, A[o + 6]);
        A[o + 7] = fma(b2, b2, A[o + 7]);
    } else if constexpr (K == 3) {
        const double u = p * q, w1 = f ? p : q, w2 = f ? q : p;
        const double L = fma(u, fma(w1, c.x3, fma(w2, fma(w1, c.x8, c.x5), c.x1)), eps);
        P = P * L;
        const double r = recip_rn(L);
        double d[4];
        d[0] = u * r;
        d[1] = d[0] * w1;
        d[2] = d[0] * w2;
        d[3] = d[1] * w2;
#pragma unroll
        for (int i = 0; i < 4; i++) A[28 + i] = A[28 + i] + d[i];
        int id = 32;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = i; j < 4; j++, id++) A[id] = fma(d[i], d[j], A[id]);
    } else {
        const double u = p * q;
        const double L = fma(u, fma(u, c.E8, c.E67), eps);
        P = P * L;
        const double r = recip_rn(L);
        const double e0 = u * r, e1 = e0 * u;
        A[42] = A[42] + e0;
        A[43] = A[43] + e1;
        A[44] = fma(e0, e0, A[44]);
        A[45] = fma(e0, e1, A[45]);
        A[46] = fma(e1, e1, A[46]);
    }
}

// position -> HS group; (factor, exponent) in B, C and E; index in D
constexpr int HS_GRP[9] = {0, 1, 1, 2, 1, 2, 1, 2, 3};
constexpr int B_F[9] = {0, 0, 1, 2, 0, 0, 0, 1, 2}, B_E[9] = {0, 0, 0, 1, 0, 0, 0, 1, 2};
constexpr int C_F[9] = {0, 0, 0, 0, 1, 2, 0, 1, 2}, C_E[9] = {0, 0, 0, 0, 0, 1, 0, 1, 2};
constexpr int D_I[9] = {-1, 0, -1, 1, -1, 2, -1, -1, 3};
constexpr int E_F[9] = {0, 0, 0, 0, 0, 0, 2, 1, 4}, E_E[9] = {0, 0, 0, 0, 0, 0, 0, 0, 1};
constexpr int TRI4_BASE[4] = {0, 4, 7, 9};

// the 47 sums -> g[9], h[45] (the oracle's ko_class_expand), each value
// stored as it is formed: out[k * stride] for k = 1 + i (g) and 10 + id (h)
__device__ __forceinline__ void class_expand_store(const double* R, double* out, size_t stride) {
    const double *gA = R + 1, *SA = R + 5, *gB = R + 12, *SB = R + 15, *gC = R + 20, *SC = R + 23;
    const double *gD = R + 28, *SD = R + 32, *gE = R + 42, *SE = R + 44;
    out[1 * stride] = gA[0];
    out[2 * stride] = gA[1] + gD[0];
    out[3 * stride] = gA[1] + gB[0];
    out[4 * stride] = (gA[2] + 2.0 * gB[1]) + gD[1];
    out[5 * stride] = gA[1] + gC[0];
    out[6 * stride] = (gA[2] + 2.0 * gC[1]) + gD[2];
    out[7 * stride] = gA[1] + 2.0 * gE[0];
    out[8 * stride] = ((gA[2] + gB[1]) + gC[1]) + gE[0];
    out[9 * stride] = (((gA[3] + 2.0 * gB[2]) + 2.0 * gC[2]) + gD[3]) + 4.0 * gE[1];
    int id = 0;
#pragma unroll
    for (int a = 0; a < 9; a++)
#pragma unroll
        for (int b = a; b < 9; b++, id++) {
            double h = SA[HS_GRP[a] + HS_GRP[b]];
            if (B_F[a] && B_F[b]) h = h + (double)(B_F[a] * B_F[b]) * SB[B_E[a] + B_E[b]];
            if (C_F[a] && C_F[b]) h = h + (double)(C_F[a] * C_F[b]) * SC[C_E[a] + C_E[b]];
            if (D_I[a] >= 0 && D_I[b] >= 0) h = h + SD[TRI4_BASE[D_I[a]] + (D_I[b] - D_I[a])];
            if (E_F[a] && E_F[b]) h = h + (double)(E_F[a] * E_F[b]) * SE[E_E[a] + E_E[b]];
            out[(size_t)(10 + id) * stride] = h;
        }
}

// the chunks of one pattern's list in a block (entries blk[off .. off + n)):
// each lane's entry is loaded three chunks ahead and its frequencies read from
// LDS one chunk ahead (reads past the list stay inside the block's region and
// are masked to a site of the block; their lanes compute nothing); t counts
// the segment's chunks, and every 16th ends each lane's log group
template <int K>
__device__ __forceinline__ void class_pattern(const uint16_t* __restrict__ blk, uint32_t off, uint32_t n,
                                              uint32_t lane, const double2* afb, double* A, double& P, uint32_t& t,
                                              const ClassCoef& c, double eps) {
    if (n == 0) return;
    auto entry = [&](uint32_t j) { return (uint32_t)blk[min(off + j, (uint32_t)CLASS_BLOCK - 1)]; };
    uint32_t e = entry(lane), e1 = entry(64 + lane), e2 = entry(128 + lane);
    double2 f = afb[(e >> 1) & (CLASS_BLOCK - 1)];
    for (uint32_t j0 = 0; j0 < n; j0 += 64) {
        const uint32_t e3 = entry(j0 + 192 + lane);
        const double2 f1 = afb[(e1 >> 1) & (CLASS_BLOCK - 1)];
        if (j0 + lane < n) class_site<K>(A, P, e & 1u, f.x, f.y, c, eps);
        if ((++t & (LOG_GROUP - 1)) == 0) {
            A[0] = A[0] + sleef_log_pos(P);
            P = 1.0;
        }
        e = e1;
        e1 = e2;
        e2 = e3;
        f = f1;
    }
}

}  // namespace

// W wavefronts a workgroup, one slot each; blockIdx.y = segment, whose bounds
// are multiples of CLASS_BLOCK (the last ends the rows).  Per block the
// workgroup stages the sites' (p, q) in LDS, then each wavefront walks its
// pair's five lists.  Partials (f, g[9], h[45], n) as sweep_fp64 writes them.
template <int W>
__global__ __launch_bounds__(W * 64, 3) void sweep_class_fp64(
    const double* __restrict__ af, uint32_t v_pad, const uint32_t* __restrict__ seg_bounds,
    const uint16_t* __restrict__ lists, const uint16_t* __restrict__ counts, uint32_t nblk,
    const double* __restrict__ pt, const int* __restrict__ active, uint32_t nslots, double eps,
    double* __restrict__ partial) {
    __shared__ double2 afb[CLASS_BLOCK];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t slot = blockIdx.x * W + w;
    const bool live = slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;
    const uint32_t seg = blockIdx.y;
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    const uint32_t bb0 = s0 / CLASS_BLOCK, bb1 = (s1 + CLASS_BLOCK - 1) / CLASS_BLOCK;

    double D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) D[k] = live ? pt[(size_t)k * nslots + slot] : 0.0;
    const ClassCoef c = class_coef(D);
    double A[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) A[k] = 0.0;
    double P = 1.0;
    uint32_t t = 0, n = 0;
    for (uint32_t b = bb0; b < bb1; b++) {
        const uint32_t site0 = b * CLASS_BLOCK;
        const uint32_t nsite = min((uint32_t)CLASS_BLOCK, v_pad - site0);
        __syncthreads();  // the previous block's frequencies are no longer read
        for (uint32_t i = threadIdx.x; i < nsite; i += W * 64)
            afb[i] = *reinterpret_cast<const double2*>(af + 2 * ((size_t)site0 + i));
        __syncthreads();
        if (!live) continue;
        const uint4 cv = reinterpret_cast<const uint4*>(counts)[(size_t)slot * nblk + b];
        const uint32_t nk[5] = {cv.x & 0xffffu, cv.x >> 16, cv.y & 0xffffu, cv.y >> 16, cv.z};
        const uint16_t* blk = lists + ((size_t)slot * nblk + b) * CLASS_BLOCK;
        uint32_t off = 0;
        class_pattern<0>(blk, off, nk[0], lane, afb, A, P, t, c, eps);
        off += nk[0];
        class_pattern<1>(blk, off, nk[1], lane, afb, A, P, t, c, eps);
        off += nk[1];
        class_pattern<2>(blk, off, nk[2], lane, afb, A, P, t, c, eps);
        off += nk[2];
        class_pattern<3>(blk, off, nk[3], lane, afb, A, P, t, c, eps);
        off += nk[3];
        class_pattern<4>(blk, off, nk[4], lane, afb, A, P, t, c, eps);
        n += nk[0] + nk[1] + nk[2] + nk[3] + nk[4];
    }
    if (!live) return;
    if (t & (LOG_GROUP - 1)) A[0] = A[0] + sleef_log_pos(P);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < NACC; k++) A[k] = A[k] + __shfl_xor(A[k], o, 64);
    if (lane == 0) {
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = A[0];
        class_expand_store(A, out, nslots);
        out[(size_t)55 * nslots] = (double)n;
    }
}

namespace klaunch {

hipError_t class_build(const uint32_t* geno, uint32_t row_words, const uint32_t* sx, const uint32_t* sy,
                       const int* active, uint32_t nslots, uint32_t nblk, uint16_t* lists, uint16_t* counts,
                       uint2* built, hipStream_t st) {
    hipLaunchKernelGGL(class_build_kernel, dim3((nslots + 3) / 4), dim3(256), 0, st, geno, row_words, sx, sy, active,
                       nslots, nblk, lists, counts, built);
    return hipGetLastError();
}

hipError_t sweep_class(const double* af, uint32_t v_pad, const uint32_t* seg_bounds, uint32_t nseg,
                       const uint16_t* lists, const uint16_t* counts, uint32_t nblk, const double* pt,
                       const int* active, uint32_t nslots, double eps, double* partial, hipStream_t st, int waves) {
    auto go = [&](auto kern, int W) {
        hipLaunchKernelGGL(kern, dim3((nslots + W - 1) / W, nseg), dim3(W * 64), 0, st, af, v_pad, seg_bounds, lists,
                           counts, nblk, pt, active, nslots, eps, partial);
    };
    if (waves == 4)
        go(sweep_class_fp64<4>, 4);
    else if (waves == 8)
        go(sweep_class_fp64<8>, 8);
    else if (waves == 12)
        go(sweep_class_fp64<12>, 12);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace klaunch
