// hessian_forms.hip -- timing of the Hessian accumulation H = sum_s c'_s c'_s^T
// (45 unique terms, the per-site rank-1 update of jacquard/grad_hess.rs:149-448)
// in the forms a sweep could give it, on the same synthetic scaled rows:
//
//   form 0  one pair per lane, FP64 FMA (the FP64 sweeps)
//   form 1  one pair per lane, FP32 FMA (the mixed-precision sweeps)
//   form 2  one pair per wavefront, its sites across the 64 lanes; c' split
//           into bf16 hi + lo (hi: c' with its low 16 bits cleared, lo: the
//           exact remainder cut to bf16 the same way: |c' - hi| < 2^-7 |c'|,
//           |c' - hi - lo| < 2^-14 |c'|, so a product misses less than
//           3 2^-14 of itself), transposed through LDS into the operand layout of
//           v_mfma_f32_16x16x32_bf16, H = hi hi^T + hi lo^T + (hi lo^T)^T
//           accumulated by the matrix cores in FP32
//   form 3  as form 2 with a three-way split hi + mid + lo (four products:
//           hh, hm, hl, mm, the transposed ones added at the end; a product
//           misses less than 3 2^-21 of itself)
//   form 4  form 1 without the Hessian (the rows and the gradient sums only)
//   form 5  form 2 without the Hessian
//   form 6  form 2's layout with the Hessian by FP32 FMA in each lane over
//           its own sites, the lanes' sums added across the wavefront at the
//           segment's end
//
// Every form builds the same rows: c'_k = T[s][k] * (T[s][9] * w), FP32
// products of a 256-site table T and a per-pair scale w = 1 + (pair % 64) / 64,
// the shape of the mixed sweep's c' = c * (1/L); and every form sums the
// gradient terms g_k += c'_k in FP32.  As in the sweeps, the sites are cut
// into segments of SEG sites (blockIdx.y), each summed on its own.  Outputs,
// per segment and pair: 45 Hessian terms (lower triangle, row by row) then 9
// gradient sums.  This is measurement code: nothing in a fit calls it.
#include <hip/hip_runtime.h>

#include "host.hpp"

namespace {

constexpr int HT = 256;   // table sites
constexpr int HW = 12;    // floats a table row: 9 c, the row scale, 2 of padding
constexpr int TRS = 72;   // bf16 a transposed row: 64 sites and 8 of padding
constexpr int NOUT = 54;  // 45 Hessian terms, 9 gradient sums
constexpr uint32_t SEG = 4096;  // sites a segment

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS written by a wavefront's lanes, then read by others of the same
// wavefront: its LDS operations complete in order, so a fence at wavefront
// scope (which keeps the compiler from moving them) is all it takes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void load_table(const float* __restrict__ table, float* tab) {
    for (int i = threadIdx.x; i < HT * HW; i += blockDim.x) tab[i] = table[i];
}

__device__ __forceinline__ void scaled_row(const float* tab, uint32_t s, float w, float* c) {
    const float* t = tab + (s & (HT - 1)) * HW;
    const float r = t[9] * w;
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = t[k] * r;
}

// forms 0, 1, 4: one pair per lane, every lane walks all sites
template <typename T, bool WITH_H>
__global__ __launch_bounds__(256) void hess_lane_kernel(const float* __restrict__ table, uint32_t n_pairs,
                                                        uint32_t n_sites, T* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) float tab[HT * HW];
    load_table(table, tab);
    __syncthreads();
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const float w = 1.0f + (float)(p & 63u) / 64.0f;
    const uint32_t s_begin = blockIdx.y * SEG, s_end = s_begin + SEG;
    T H[45], G[9];
#pragma unroll
    for (int k = 0; k < 45; k++) H[k] = T(0);
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = T(0);
#pragma unroll 4
    for (uint32_t s = s_begin; s < s_end; s++) {
        float c[9];
        scaled_row(tab, s, w, c);
#pragma unroll
        for (int k = 0; k < 9; k++) G[k] += (T)c[k];
        if (WITH_H) {
            int q = 0;
#pragma unroll
            for (int i = 0; i < 9; i++)
#pragma unroll
                for (int j = 0; j <= i; j++, q++) H[q] = fma((T)c[i], (T)c[j], H[q]);
        }
    }
    if (p < n_pairs) {
        T* po = out + ((size_t)blockIdx.y * n_pairs + p) * NOUT;
#pragma unroll
        for (int k = 0; k < 45; k++) po[k] = H[k];
#pragma unroll
        for (int k = 0; k < 9; k++) po[45 + k] = G[k];
    }
}

// forms 2, 3, 5, 6: one pair per wavefront (four a workgroup), lane l takes
// the sites s0 + l of each 64-site step.  NPART = 2 (hi, lo) or 3 (hi, mid,
// lo); NPART = 0: no Hessian; NPART = -1: FP32 FMA per lane (form 6).
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (A 16 x 32, B 32 x 16, D 16 x 16):
// lane m holds A[m % 16][8 (m / 16) + i] and B[8 (m / 16) + i][m % 16] for
// i = 0..7, and D[4 (m / 16) + i][m % 16] for i = 0..3.  With rows of A = the
// 9 components (7 rows of zeros) and K = 32 sites, A = C'^T and B = C' hold
// the same values in every lane, so one register serves both operands.
template <int NPART>
__global__ __launch_bounds__(256) void hess_mfma_kernel(const float* __restrict__ table, uint32_t n_pairs,
                                                        uint32_t n_sites, float* __restrict__ out) {
    constexpr int NP = NPART > 0 ? NPART : 1;
    __shared__ __attribute__((aligned(16))) float tab[HT * HW];
    __shared__ __attribute__((aligned(16))) unsigned short tr[4][NP][16 * TRS];  // bf16 bits
    const uint32_t l = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    load_table(table, tab);
    for (int i = l; i < NP * 16 * TRS; i += 64) (&tr[wv][0][0])[i] = 0;
    __syncthreads();
    const uint32_t p = blockIdx.x * 4 + wv;
    const float w = 1.0f + (float)(p & 63u) / 64.0f;
    float G[9], HL[NPART < 0 ? 45 : 1];
#pragma unroll
    for (int k = 0; k < 9; k++) G[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < (NPART < 0 ? 45 : 1); k++) HL[k] = 0.0f;
    // accumulators: hh, h*mid or h*lo, ... (see the header)
    // (two sets, one per 32-site half of a step: independent chains)
    constexpr int NACC = NPART == 3 ? 4 : NPART == 2 ? 2 : 0;
    f32x4 acc[2][NACC > 0 ? NACC : 1];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int a = 0; a < (NACC > 0 ? NACC : 1); a++) acc[h][a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int row = l & 15, kb = 8 * (l >> 4);
    for (uint32_t s0 = blockIdx.y * SEG; s0 < (blockIdx.y + 1) * SEG; s0 += 64) {
        float c[9];
        scaled_row(tab, s0 + l, w, c);
#pragma unroll
        for (int k = 0; k < 9; k++) G[k] += c[k];
        if (NPART < 0) {
            int q = 0;
#pragma unroll
            for (int i = 0; i < 9; i++)
#pragma unroll
                for (int j = 0; j <= i; j++, q++) HL[q] = fmaf(c[i], c[j], HL[q]);
        }
        if (NPART > 0) {
#pragma unroll
            for (int k = 0; k < 9; k++) {
                float rest = c[k];
#pragma unroll
                for (int t = 0; t < NP; t++) {
                    const uint32_t b = __float_as_uint(rest);
                    tr[wv][t][k * TRS + l] = (unsigned short)(b >> 16);
                    if (t + 1 < NP) rest = rest - __uint_as_float(b & 0xffff0000u);  // exact
                }
            }
            wave_sync();  // the wavefront's rows are written (its own buffer)
#pragma unroll
            for (int half = 0; half < 2; half++) {
                bf16x8 v[NP];
#pragma unroll
                for (int t = 0; t < NP; t++)
                    v[t] = __builtin_bit_cast(bf16x8,
                                              *reinterpret_cast<const u16x8*>(&tr[wv][t][row * TRS + 32 * half + kb]));
                f32x4* ac = acc[half];
                ac[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v[0], v[0], ac[0], 0, 0, 0);
                if (NPART == 2) {
                    ac[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v[0], v[NP - 1], ac[1], 0, 0, 0);
                } else if (NPART == 3) {
                    ac[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v[0], v[1], ac[1], 0, 0, 0);
                    ac[2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v[0], v[NP - 1], ac[2], 0, 0, 0);
                    ac[3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v[1], v[1], ac[3], 0, 0, 0);
                }
            }
            wave_sync();  // the operands are read before the next step writes
        }
    }
    // the gradient sums over the wavefront's lanes
#pragma unroll
    for (int k = 0; k < 9; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) G[k] += __shfl_xor(G[k], o, 64);
    float* po = out + ((size_t)blockIdx.y * n_pairs + p) * NOUT;
    if (NPART < 0) {
#pragma unroll
        for (int k = 0; k < 45; k++)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) HL[k] += __shfl_xor(HL[k], o, 64);
        if (l == 0 && p < n_pairs)
#pragma unroll
            for (int k = 0; k < 45; k++) po[k] = HL[k];
    } else if (NPART > 0) {
        // the wavefront's accumulators as 16 x 16 matrices, in its own
        // transpose buffer (NACC x 1 KB within NP x 2.3 KB; the last step's
        // operands were read before the loop's closing barrier)
        float* dsum = reinterpret_cast<float*>(&tr[wv][0][0]);
#pragma unroll
        for (int a = 0; a < NACC; a++)
#pragma unroll
            for (int i = 0; i < 4; i++) dsum[a * 256 + (4 * (l >> 4) + i) * 16 + (l & 15)] = acc[0][a][i] + acc[1][a][i];
        wave_sync();
        if (l < 45 && p < n_pairs) {
            int i = 0;
            while ((i + 1) * (i + 2) / 2 <= (int)l) i++;
            const int j = (int)l - i * (i + 1) / 2;
            const float* d0 = dsum;
            float h = d0[i * 16 + j];
            // the cross products and their transposes
            for (int a = 1; a < NACC; a++) {
                const float* d = dsum + a * 256;
                if (NPART == 3 && a == 3)
                    h += d[i * 16 + j];  // mid mid^T is symmetric already
                else
                    h += d[i * 16 + j] + d[j * 16 + i];
            }
            po[l] = h;
        }
    } else if (l < 45 && p < n_pairs) {
        po[l] = 0.0f;
    }
    if (l == 0 && p < n_pairs)
#pragma unroll
        for (int k = 0; k < 9; k++) po[45 + k] = G[k];
}

}  // namespace

extern "C" int kestrel_measure_hessian_form(int form, uint32_t n_pairs, uint32_t n_sites, const float* table,
                                            void* out, double* ms) {
    if (!table || !out || !ms) return fail(KESTREL_EINVAL, "null argument");
    if (form < 0 || form > 6) return fail(KESTREL_EINVAL, "form %d is not 0..6", form);
    if (n_pairs == 0 || n_pairs % 256 || n_sites == 0 || n_sites % SEG)
        return fail(KESTREL_EINVAL, "n_pairs must be a positive multiple of 256 and n_sites of %u", SEG);
    const uint32_t nseg = n_sites / SEG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KESTREL_EDEVICE, "no HIP device: libkestrel_amd needs an MI355X (gfx950) GPU");
    const size_t esz = form == 0 ? sizeof(double) : sizeof(float);
    DevBuf<float> dtab;
    DevBuf<unsigned char> dout;
    HIPCHK(dtab.alloc((size_t)HT * HW));
    HIPCHK(dout.alloc((size_t)nseg * n_pairs * NOUT * esz));
    HIPCHK(hipMemcpy(dtab.p, table, sizeof(float) * HT * HW, hipMemcpyHostToDevice));
    auto launch = [&]() {
        float* o32 = reinterpret_cast<float*>(dout.p);
        switch (form) {
            case 0:
                hipLaunchKernelGGL((hess_lane_kernel<double, true>), dim3(n_pairs / 256, nseg), dim3(256), 0, nullptr,
                                   dtab.p, n_pairs, n_sites, reinterpret_cast<double*>(dout.p));
                break;
            case 1:
                hipLaunchKernelGGL((hess_lane_kernel<float, true>), dim3(n_pairs / 256, nseg), dim3(256), 0, nullptr,
                                   dtab.p, n_pairs, n_sites, o32);
                break;
            case 2:
                hipLaunchKernelGGL(hess_mfma_kernel<2>, dim3(n_pairs / 4, nseg), dim3(256), 0, nullptr, dtab.p, n_pairs,
                                   n_sites, o32);
                break;
            case 3:
                hipLaunchKernelGGL(hess_mfma_kernel<3>, dim3(n_pairs / 4, nseg), dim3(256), 0, nullptr, dtab.p, n_pairs,
                                   n_sites, o32);
                break;
            case 4:
                hipLaunchKernelGGL((hess_lane_kernel<float, false>), dim3(n_pairs / 256, nseg), dim3(256), 0, nullptr,
                                   dtab.p, n_pairs, n_sites, o32);
                break;
            case 6:
                hipLaunchKernelGGL(hess_mfma_kernel<-1>, dim3(n_pairs / 4, nseg), dim3(256), 0, nullptr, dtab.p,
                                   n_pairs, n_sites, o32);
                break;
            default:
                hipLaunchKernelGGL(hess_mfma_kernel<0>, dim3(n_pairs / 4, nseg), dim3(256), 0, nullptr, dtab.p, n_pairs,
                                   n_sites, o32);
        }
        return hipGetLastError();
    };
    struct Events {  // destroyed on every return path
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    HIPCHK(hipEventCreate(&ev.e[0]));
    HIPCHK(hipEventCreate(&ev.e[1]));
    HIPCHK(launch());  // warm up
    HIPCHK(hipEventRecord(ev.e[0], nullptr));
    HIPCHK(launch());
    HIPCHK(hipEventRecord(ev.e[1], nullptr));
    HIPCHK(hipEventSynchronize(ev.e[1]));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ev.e[0], ev.e[1]));
    *ms = t;
    HIPCHK(hipMemcpy(out, dout.p, (size_t)nseg * n_pairs * NOUT * esz, hipMemcpyDeviceToHost));
    return KESTREL_OK;
}
