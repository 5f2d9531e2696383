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
// genotype pair) and the five list lengths.  Entry j of a pattern's list is
// lane j % 64's site in chunk j / 64; in memory each pattern starts on a group
// of 256 entries, and a group holds lane l's entries of its four chunks side
// by side (entry j at (j / 256) 256 + 4 (j % 64) + (j / 64) % 4), so one
// 8-byte load gives a lane four chunks.  sweep_class_fp64 stages each block's
// frequencies in LDS for the workgroup's pairs (global_load_lds, no registers)
// and walks the groups of the block's lists, loaded two groups ahead; an
// entry addresses the frequency its flag selects directly.
//
// The summation order (each lane's sums in list order, a log of the lane's
// product after the first group that ends at least 13 chunks after the last
// log and at the segment's end, a butterfly over the 64 lanes, the expansion)
// is restated by the oracle's class order;
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

// One workgroup per slot whose pair (sx, sy) differs from the one its lists
// were built for (built[slot]); its BUILD_WAVES wavefronts sort one block at
// a time.  Thread t takes the sites BUILD_SITES t .. BUILD_SITES (t + 1) - 1
// of the block (BUILD_WORDS words of each row, loaded a block ahead), counts
// its sites of each pattern, and a workgroup scan (wavefront scans, then the
// wavefront totals through LDS) gives each thread the position of its first
// site in its pattern's list; the thread then writes its entries in site
// order into the LDS copy of the block's lists through per-thread LDS
// counters (pattern k's list starting at the group-aligned pb[k]).  The copy
// is written out in the group layout: lane l's entries j, j + 64, j + 128,
// j + 192 of a group side by side.
constexpr int BUILD_WAVES = 4;
constexpr int BUILD_THREADS = BUILD_WAVES * 64;
constexpr int BUILD_SITES = CLASS_BLOCK / BUILD_THREADS;  // sites of a block a thread takes
constexpr int BUILD_WORDS = BUILD_SITES / 16;             // 2-bit words of a row a thread takes
constexpr uint32_t MISS = 0xffffu;  // above every entry ((site << 1) | flag < 2 CLASS_BLOCK)
// a block's lists fill at most CLASS_BLOCK entries + 4 partial groups
static_assert(2 * CLASS_BLOCK <= (int)MISS && CLASS_BLOCK + 4 * CLASS_GROUP < CLASS_STRIDE, "stage bounds");
static_assert(BUILD_WORDS == 2, "a thread loads its words of a row as one uint2");

__global__ __launch_bounds__(BUILD_THREADS) void class_build_kernel(
    const uint32_t* __restrict__ geno, uint32_t row_words, const uint32_t* __restrict__ sx,
    const uint32_t* __restrict__ sy, const int* __restrict__ active, uint32_t nslots, uint32_t slot0, uint32_t nchunk,
    uint32_t nblk, uint16_t* __restrict__ lists, uint16_t* __restrict__ counts, uint2* __restrict__ built) {
    __shared__ uint32_t ctr[5 * BUILD_THREADS];     // pattern k's counter of thread t at k BUILD_THREADS + t
    __shared__ uint32_t wtot[BUILD_WAVES][4];       // each wavefront's packed pattern counts
    __shared__ __attribute__((aligned(16))) uint16_t stage[CLASS_STRIDE];
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(t >> 6));
    const uint32_t ls = blockIdx.x, slot = slot0 + ls;  // the lists' slot (in the chunk), the slot
    if (ls >= nchunk || slot >= nslots || active[slot] == 0) return;  // uniform over the workgroup
    const uint32_t x = sx[slot], y = sy[slot];
    const uint2 b0 = built[ls];
    if (b0.x == x && b0.y == y) return;
    const uint32_t* rowx = geno + (size_t)x * row_words;
    const uint32_t* rowy = geno + (size_t)y * row_words;
    // thread t's words of block b (missing calls past the row's end; row_words
    // is a multiple of 16, so both words are in the row or neither)
    auto words = [&](uint32_t b, uint2& wx, uint2& wy) {
        const uint32_t w0 = b * (CLASS_BLOCK / 16) + BUILD_WORDS * t;
        wx = wy = make_uint2(0xffffffffu, 0xffffffffu);
        if (w0 < row_words) {
            wx = *reinterpret_cast<const uint2*>(rowx + w0);
            wy = *reinterpret_cast<const uint2*>(rowy + w0);
        }
    };
    uint2 nx, ny;
    words(0, nx, ny);
    for (uint32_t b = 0; b < nblk; b++) {
        const uint32_t ax[BUILD_WORDS] = {nx.x, nx.y}, ay[BUILD_WORDS] = {ny.x, ny.y};
        if (b + 1 < nblk) words(b + 1, nx, ny);
        uint32_t cnt[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < BUILD_WORDS; q++) {
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
        // counts packed two to a word (each at most CLASS_BLOCK < 2^16)
        const uint32_t s01 = wave_scan(cnt[0] | (cnt[1] << 16), lane);
        const uint32_t s23 = wave_scan(cnt[2] | (cnt[3] << 16), lane);
        const uint32_t s4 = wave_scan(cnt[4], lane);
        if (lane == 63) {
            wtot[w][0] = s01;
            wtot[w][1] = s23;
            wtot[w][2] = s4;
        }
        __syncthreads();
        uint32_t o01 = 0, o23 = 0, o4 = 0, t01 = 0, t23 = 0, t4 = 0;
#pragma unroll
        for (int v = 0; v < BUILD_WAVES; v++) {
            const uint32_t a = wtot[v][0], c = wtot[v][1], e = wtot[v][2];
            if ((uint32_t)v < w) {
                o01 += a;
                o23 += c;
                o4 += e;
            }
            t01 += a;
            t23 += c;
            t4 += e;
        }
        const uint32_t tot[5] = {t01 & 0xffffu, t01 >> 16, t23 & 0xffffu, t23 >> 16, t4};
        const uint32_t i01 = s01 + o01, i23 = s23 + o23, i4 = s4 + o4;  // inclusive over the workgroup
        const uint32_t inc[5] = {i01 & 0xffffu, i01 >> 16, i23 & 0xffffu, i23 >> 16, i4};
        uint32_t pb[6];
        pb[0] = 0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            ctr[k * BUILD_THREADS + t] = pb[k] + inc[k] - cnt[k];
            pb[k + 1] = pb[k] + (tot[k] + CLASS_GROUP - 1) / CLASS_GROUP * CLASS_GROUP;
        }
        // a word's 16 sites: the 16 counter increments issued back to back
        // (a thread's increments of one counter return in issue order, so its
        // entries keep site order), then the 16 entries written; a missing
        // site (entry MISS) adds 0 and writes the stage's last entry, which
        // no list reaches
#pragma unroll
        for (int q = 0; q < BUILD_WORDS; q++) {
            uint32_t pos[16], ent[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t idx = (((ax[q] >> (2 * i)) & 3u) << 2) | ((ay[q] >> (2 * i)) & 3u);
                const uint32_t nib = (uint32_t)(PATTERN_LUT >> (4 * idx)) & 15u;
                const uint32_t k = nib & 7u;
                const bool miss = k == 7u;
                pos[i] = atomicAdd(&ctr[(miss ? 0u : k) * BUILD_THREADS + t], miss ? 0u : 1u);
                ent[i] = miss ? MISS : ((BUILD_SITES * t + 16u * q + (uint32_t)i) << 1) | (nib >> 3);
            }
#pragma unroll
            for (int i = 0; i < 16; i++) stage[ent[i] == MISS ? CLASS_STRIDE - 1 : pos[i]] = (uint16_t)ent[i];
        }
        __syncthreads();  // the block's lists complete in LDS
        // out in the group layout: a lane's four entries of a group side by side
        uint2* dst = reinterpret_cast<uint2*>(lists + ((size_t)ls * nblk + b) * CLASS_STRIDE);
        for (uint32_t g = w; g < pb[5] / CLASS_GROUP; g += BUILD_WAVES) {
            const uint16_t* src = stage + g * CLASS_GROUP + lane;
            dst[(size_t)g * 64 + lane] =
                make_uint2((uint32_t)src[0] | ((uint32_t)src[64] << 16), (uint32_t)src[128] | ((uint32_t)src[192] << 16));
        }
        if (t == 0) {
            uint4 cv;
            cv.x = tot[0] | (tot[1] << 16);
            cv.y = tot[2] | (tot[3] << 16);
            cv.z = tot[4];
            cv.w = 0;
            reinterpret_cast<uint4*>(counts)[(size_t)ls * nblk + b] = cv;
        }
        __syncthreads();  // the stage and the wavefront totals are free for the next block
    }
    if (t == 0) built[ls] = make_uint2(x, y);
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
// class order).  (v0, v1) are the site's frequencies as class_freqs reads
// them: v0 = the one the entry's flag selects, v1 = the other (E: p, q).
template <int K>
__device__ __forceinline__ void class_site(double* A, double& P, double v0, double v1, const ClassCoef& c,
                                           double eps) {
    if constexpr (K == 0) {
        const double a = v0;
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
        const double u = v0 * v1, w = v0;  // (p q = q p bit for bit)
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
        const double u = v1 * v0, w1 = v1, w2 = v0;
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
        const double u = v0 * v1;
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

// One group of pattern K's list: the four chunks j0 = jg, jg + 64, ... whose
// entries are q (lane l's four entries side by side); a chunk past the list's
// end stops the group.  t counts the segment's chunks, and every 16th ends
// each lane's log group.
// the log of each lane's product once a group ends at least LOG_MIN chunks
// after the last log (a group adds at most four: a product never has more
// than 16 factors)
constexpr uint32_t LOG_MIN = LOG_GROUP - 3;

__device__ __forceinline__ void class_log(double* A, double& P, uint32_t& since) {
    if (since >= LOG_MIN) {
        A[0] = A[0] + sleef_log_pos(P);
        P = 1.0;
        since = 0;
    }
}

// The frequencies of list entry e = (site << 1) | flag from the block's
// (p, q) pairs in LDS, read as doubles: afd[e] is p for flag 0 and q for
// flag 1, afd[e ^ 1] the other, so the entry selects without a compare.
// E sites (flag 0, e even) read the pair whole.  Entries within a list are below
// 2 CLASS_BLOCK; MASK keeps the lanes past a list's end (whose entries are
// not written) inside the block's frequencies.
template <int K, bool MASK>
__device__ __forceinline__ void class_freqs(const double* afd, uint32_t e, double& v0, double& v1) {
    const uint32_t i = MASK ? e & (2 * CLASS_BLOCK - 1) : e;
    const uint32_t o = i << 3;  // byte offsets: the partner is 8 bytes away
    const char* base = reinterpret_cast<const char*>(afd);
    if constexpr (K == 4) {
        const double2 f = *reinterpret_cast<const double2*>(base + o);
        v0 = f.x;
        v1 = f.y;
    } else {
        v0 = *reinterpret_cast<const double*>(base + o);
        v1 = K == 0 ? 0.0 : *reinterpret_cast<const double*>(base + (o ^ 8u));
    }
}

// One group of pattern K's list: the chunks j0 = jg, jg + 64, jg + 128,
// jg + 192 below n, whose entries are q (lane l's four entries side by side).
template <int K>
__device__ __forceinline__ void class_group(uint2 q, uint32_t jg, uint32_t n, uint32_t lane, const double* afd,
                                            double* A, double& P, uint32_t& since, const ClassCoef& c, double eps) {
    const uint32_t e0 = q.x & 0xffffu, e1 = q.x >> 16, e2 = q.y & 0xffffu, e3 = q.y >> 16;
    if (jg + CLASS_GROUP <= n) {  // four whole chunks: their frequencies read first
        double a0, b0, a1, b1, a2, b2, a3, b3;
        class_freqs<K, false>(afd, e0, a0, b0);
        class_freqs<K, false>(afd, e1, a1, b1);
        class_freqs<K, false>(afd, e2, a2, b2);
        class_freqs<K, false>(afd, e3, a3, b3);
        class_site<K>(A, P, a0, b0, c, eps);
        class_site<K>(A, P, a1, b1, c, eps);
        class_site<K>(A, P, a2, b2, c, eps);
        class_site<K>(A, P, a3, b3, c, eps);
        since += 4;
    } else {  // the list's last group: its 1 to 4 chunks, the last one partial
        const uint32_t m = (n - jg + 63) / 64;
        double a[4], b[4];
        class_freqs<K, true>(afd, e0, a[0], b[0]);
        class_freqs<K, true>(afd, e1, a[1], b[1]);
        class_freqs<K, true>(afd, e2, a[2], b[2]);
        class_freqs<K, true>(afd, e3, a[3], b[3]);
#pragma unroll
        for (uint32_t sub = 0; sub < 4; sub++)
            if (jg + 64 * sub + lane < n) class_site<K>(A, P, a[sub], b[sub], c, eps);
        since += m;
    }
    class_log(A, P, since);
}

// the (p, q) of the block's sites into LDS without registers: each 64-lane
// global_load_lds_dwordx4 moves 64 consecutive sites (a block holds a multiple
// of 64 sites: rows are padded to 256)
template <int W>
__device__ __forceinline__ void stage_af(const double* __restrict__ af, uint32_t site0, uint32_t nsite, double2* buf,
                                         uint32_t w, uint32_t lane) {
    for (uint32_t u = w; 64u * u < nsite; u += W)
        __builtin_amdgcn_global_load_lds((const void*)(af + 2 * ((size_t)site0 + 64u * u + lane)),
                                         (__attribute__((address_space(3))) void*)(buf + 64u * u), 16, 0, 0);
}

}  // namespace

// The groups [g0, g1) of one block's lists for pass PASS (0: the HS list;
// 1: the B, C, D and E lists), each lane's 8-byte share of a group loaded two
// groups ahead.
// group g (of the block's groups [g0, g1) for patterns K0..K1) of the pattern
// it belongs to; gend[k] = the groups of patterns 0..k
template <int K0, int K1>
__device__ __forceinline__ void class_group_of(uint2 q, uint32_t g, const uint32_t* nk, const uint32_t* gend,
                                               uint32_t lane, const double* afd, double* A, double& P, uint32_t& since,
                                               const ClassCoef& c, double eps) {
    if constexpr (K0 == K1) {
        class_group<K0>(q, (g - (K0 ? gend[K0 - 1] : 0)) * CLASS_GROUP, nk[K0], lane, afd, A, P, since, c, eps);
    } else {
        if (g < gend[K0])
            class_group<K0>(q, (g - gend[K0 - 1]) * CLASS_GROUP, nk[K0], lane, afd, A, P, since, c, eps);
        else
            class_group<K1>(q, (g - gend[K0]) * CLASS_GROUP, nk[K1], lane, afd, A, P, since, c, eps);
    }
}

// One pass over the lists of patterns K0..K1: per block of the segment the
// workgroup stages the block's (p, q) in LDS, then each live wavefront walks
// its pair's lists of those patterns in the block, a group (four chunks) at a
// time, two groups in flight (alternate registers, so that no loaded register
// is copied before its use).  The patterns' sums are then added over the
// lanes and kept in LDS.
template <int K0, int K1, int W>
__device__ __forceinline__ void class_pass(const double* __restrict__ af, uint32_t v_pad, uint32_t bb0, uint32_t nb,
                                           const uint16_t* __restrict__ lists, const uint16_t* __restrict__ counts,
                                           uint32_t nblk, uint32_t slot, bool live, uint32_t w, uint32_t lane,
                                           double2* afb, double* sums, double& F, double& P, uint32_t& since,
                                           uint32_t& n, const ClassCoef& c, double eps) {
    constexpr int FIRST[6] = {1, 12, 20, 28, 42, 47};  // the sums of pattern k: A[FIRST[k] .. FIRST[k + 1])
    constexpr int O = FIRST[K0], E = FIRST[K1 + 1];
    const double* afd = reinterpret_cast<const double*>(afb);
    const uint4* counts4 = reinterpret_cast<const uint4*>(counts);
    double A[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) A[k] = 0.0;
    A[0] = F;
    // block b's first two groups of the pass's lists (counts cb), loaded
    // before the block's frequencies are staged, so that their latency
    // overlaps the staging and the barrier
    uint4 cv = make_uint4(0, 0, 0, 0);
    uint2 qa = make_uint2(0, 0), qb = make_uint2(0, 0);
    auto first_groups = [&](uint32_t b, uint4 cb) {
        uint32_t g0 = 0, g1 = 0, gt = 0;
        const uint32_t nk[5] = {cb.x & 0xffffu, cb.x >> 16, cb.y & 0xffffu, cb.y >> 16, cb.z};
#pragma unroll
        for (int k = 0; k < 5; k++) {
            if (k == K0) g0 = gt;
            gt += (nk[k] + CLASS_GROUP - 1) / CLASS_GROUP;
            if (k == K1) g1 = gt;
        }
        const uint2* gp = reinterpret_cast<const uint2*>(lists + ((size_t)slot * nblk + b) * CLASS_STRIDE) + lane;
        qa = gp[(size_t)(g0 < g1 ? g0 : 0) * 64];  // (group 0 of the block: a load within the lists)
        qb = gp[(size_t)(g0 + 1 < g1 ? g0 + 1 : 0) * 64];
    };
    if (live && nb) {
        cv = counts4[(size_t)slot * nblk + bb0];
        first_groups(bb0, cv);
    }
    for (uint32_t i = 0; i < nb; i++) {
        const uint32_t b = bb0 + i;
        const uint32_t site0 = b * CLASS_BLOCK;
        __syncthreads();  // the previous block's frequencies are no longer read
        stage_af<W>(af, site0, min((uint32_t)CLASS_BLOCK, v_pad - site0), afb, w, lane);
        uint4 cvn = cv;  // the next block's counts
        if (live && i + 1 < nb) cvn = counts4[(size_t)slot * nblk + b + 1];
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (!live) continue;
        const uint32_t nk[5] = {cv.x & 0xffffu, cv.x >> 16, cv.y & 0xffffu, cv.y >> 16, cv.z};
        uint32_t gend[5], gt = 0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            gt += (nk[k] + CLASS_GROUP - 1) / CLASS_GROUP;
            gend[k] = gt;
        }
        if (K0 == 0) n += nk[0] + nk[1] + nk[2] + nk[3] + nk[4];
        const uint32_t g0 = K0 ? gend[K0 - 1] : 0, g1 = gend[K1];
        const uint2* gp = reinterpret_cast<const uint2*>(lists + ((size_t)slot * nblk + b) * CLASS_STRIDE) + lane;
        // (the loads are unconditional, clamped to the last group, so that
        // each waits only for its own group's load)
        for (uint32_t g = g0; g < g1; g += 2) {
            class_group_of<K0, K1>(qa, g, nk, gend, lane, afd, A, P, since, c, eps);
            qa = gp[(size_t)min(g + 2, g1 - 1) * 64];
            if (g + 1 < g1) class_group_of<K0, K1>(qb, g + 1, nk, gend, lane, afd, A, P, since, c, eps);
            qb = gp[(size_t)min(g + 3, g1 - 1) * 64];
        }
        if (i + 1 < nb) {
            cv = cvn;
            first_groups(b + 1, cv);
        }
    }
    F = A[0];
    if (!live) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = O; k < E; k++) A[k] = A[k] + __shfl_xor(A[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = O; k < E; k++) sums[k] = A[k];
}

// wavefronts per SIMD the register budget is set for: the workgroups whose
// LDS (the block's frequencies and the lane sums) fits a CU's 160 KiB, at
// most four wavefronts a SIMD (128 VGPRs each)
template <int W>
constexpr int class_waves_per_simd() {
    constexpr int lds = CLASS_BLOCK * 16 + W * (NACC + 1) * 8;
    constexpr int v = W * ((160 * 1024) / lds) / 4;
    return v < 4 ? v : 4;
}

// W wavefronts a workgroup, one slot each; blockIdx.y = segment, whose bounds
// are multiples of CLASS_BLOCK (the last ends the rows).  Three passes over
// the segment's blocks -- the HS lists, the B and C lists, the D and E lists
// -- so that a pass holds only its patterns' sums (and the objective) in
// registers; the lane sums of each pattern are kept in LDS for the expansion.
// Partials (f, g[9], h[45], n) as sweep_fp64 writes them.
template <int W>
__global__ __launch_bounds__(W * 64, class_waves_per_simd<W>()) void sweep_class_fp64(
    const double* __restrict__ af, uint32_t v_pad, const uint32_t* __restrict__ seg_bounds,
    const uint16_t* __restrict__ lists, const uint16_t* __restrict__ counts, uint32_t nblk,
    const double* __restrict__ pt, const int* __restrict__ active, uint32_t nslots, uint32_t slot0, uint32_t nchunk,
    double eps, double* __restrict__ partial) {
    __shared__ double2 afb[CLASS_BLOCK];
    __shared__ double sums[W][NACC + 1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t ls = blockIdx.x * W + w, slot = slot0 + ls;  // the lists' slot (in the chunk), the slot
    const bool live = ls < nchunk && slot < nslots && active[slot] != 0;
    if (!__syncthreads_or(live)) return;
    const uint32_t seg = blockIdx.y;
    const uint32_t s0 = seg_bounds[seg], s1 = seg_bounds[seg + 1];
    const uint32_t bb0 = s0 / CLASS_BLOCK, nb = (s1 + CLASS_BLOCK - 1) / CLASS_BLOCK - bb0;

    double D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) D[k] = live ? pt[(size_t)k * nslots + slot] : 0.0;
    const ClassCoef c = class_coef(D);
    double F = 0.0, P = 1.0;
    uint32_t since = 0, n = 0;
    double* my = sums[w];
    class_pass<0, 0, W>(af, v_pad, bb0, nb, lists, counts, nblk, ls, live, w, lane, afb, my, F, P, since, n, c, eps);
    class_pass<1, 2, W>(af, v_pad, bb0, nb, lists, counts, nblk, ls, live, w, lane, afb, my, F, P, since, n, c, eps);
    class_pass<3, 4, W>(af, v_pad, bb0, nb, lists, counts, nblk, ls, live, w, lane, afb, my, F, P, since, n, c, eps);
    if (!live) return;
    if (since) F = F + sleef_log_pos(P);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) F = F + __shfl_xor(F, o, 64);
    if (lane == 0) {
        double R[NACC];
        R[0] = F;
#pragma unroll
        for (int k = 1; k < NACC; k++) R[k] = my[k];
        double* out = partial + (size_t)seg * NPART * nslots + slot;
        out[0] = F;
        class_expand_store(R, out, nslots);
        out[(size_t)55 * nslots] = (double)n;
    }
}

namespace klaunch {

hipError_t class_build(const uint32_t* geno, uint32_t row_words, const uint32_t* sx, const uint32_t* sy,
                       const int* active, uint32_t nslots, uint32_t slot0, uint32_t nchunk, uint32_t nblk,
                       uint16_t* lists, uint16_t* counts, uint2* built, hipStream_t st) {
    if (nchunk == 0) return hipSuccess;
    hipLaunchKernelGGL(class_build_kernel, dim3(nchunk), dim3(BUILD_THREADS), 0, st, geno, row_words, sx, sy, active,
                       nslots, slot0, nchunk, nblk, lists, counts, built);
    return hipGetLastError();
}

hipError_t sweep_class(const double* af, uint32_t v_pad, const uint32_t* seg_bounds, uint32_t nseg,
                       const uint16_t* lists, const uint16_t* counts, uint32_t nblk, const double* pt,
                       const int* active, uint32_t nslots, uint32_t slot0, uint32_t nchunk, double eps, double* partial,
                       hipStream_t st, int waves) {
    if (nchunk == 0) return hipSuccess;
    auto go = [&](auto kern, int W) {
        hipLaunchKernelGGL(kern, dim3((nchunk + W - 1) / W, nseg), dim3(W * 64), 0, st, af, v_pad, seg_bounds, lists,
                           counts, nblk, pt, active, nslots, slot0, nchunk, eps, partial);
    };
    if (waves == 8)
        go(sweep_class_fp64<8>, 8);
    else if (waves == 12)
        go(sweep_class_fp64<12>, 12);
    else if (waves == 14)
        go(sweep_class_fp64<14>, 14);
    else if (waves == 16)
        go(sweep_class_fp64<16>, 16);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace klaunch
