/*
 * ko_lanes_fast.c -- the oracle's reference-order loops written in the
 * reference's own AVX-512 loop shape (test infrastructure only: the CPU
 * baseline of the benchmark times it; see kestrel_oracle.h, KO_ORDER_FAST).
 *
 * Bit-identical to the 8-lane loops of ko_lanes.c (KO_ORDER_REF): every lane
 * performs the same IEEE operations in the same order; only the loop shape
 * differs, as the reference's does:
 *   - objective (jacquard/objective.rs:22-52): per 8-site block the 9-FMA L
 *     chain, the getexp/getmant log of log/avx512.rs:15-79, a running lane
 *     sum; radd; the scalar remainder;
 *   - gradient / Hessian (jacquard/grad_hess.rs:46-449): tiles of 32 blocks,
 *     a first pass for L, 1/L, the gradient and the scaled columns c/L into
 *     a tile buffer, then the 45 Hessian sums in three register-blocked
 *     passes of 15 accumulators over the buffer, each accumulator a chain
 *     from zero over the tile added to the running sums;
 *   - c_s rows (mixture.rs:62-236 mixture_row2 on 8-site blocks) pushed
 *     row by row through the NaN filter of LaneVectorBuilder
 *     (lanevector.rs:87-139), the hard calls read as one-hot likelihoods.
 * Compiled only with AVX-512 (F/DQ) and -ffp-contract=off.
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ko_internal.h"

/* log/mod.rs:3-14 constants */
#define LOG_2_HI 0.693147180559945286226764
#define LOG_2_LO 2.319046813846299558417771e-17
#define C_6 0.1532076988502701353
#define C_5 0.1525629051003428716
#define C_4 0.1818605932937785996
#define C_3 0.2222214519839380009
#define C_2 0.2857142932794299317
#define C_1 0.399999999963525199
#define C_0 0.6666666666667333541

/* The SLEEF-u10 log of log/avx512.rs on 8 lanes: d = m 2^e with m in
 * [0.75, 1.5) by getexp / getmant, x = (m-1)/(m+1) as a double-word,
 * log = e ln2 + 2x + x^3 P(x^2); special values by masks (0 -> -inf,
 * negative or NaN -> NaN, +inf -> +inf). */
static inline __m512d log8(__m512d d) {
    const __m512d one = _mm512_set1_pd(1.0);
    __m512d e = _mm512_getexp_pd(d);
    const __m512d m = _mm512_getmant_pd(d, _MM_MANT_NORM_p75_1p5, _MM_MANT_SIGN_zero);
    e = _mm512_mask_add_pd(e, _mm512_cmp_pd_mask(m, one, _CMP_LT_OQ), e, one);
    const __m512d u = _mm512_sub_pd(m, one);
    const __m512d l0 = _mm512_add_pd(one, m);
    const __m512d lz = _mm512_sub_pd(l0, one);
    const __m512d l1 = _mm512_sub_pd(m, lz);
    const __m512d r = _mm512_div_pd(one, l0);
    const __m512d q0 = _mm512_mul_pd(u, r);
    const __m512d uu = _mm512_fmsub_pd(u, r, q0);
    const __m512d ee = _mm512_fnmadd_pd(l0, r, one);
    __m512d q1 = _mm512_fnmadd_pd(l1, r, ee);
    q1 = _mm512_fmadd_pd(q0, q1, uu);
    const __m512d x2 = _mm512_mul_pd(q0, q0);
    const __m512d x4 = _mm512_mul_pd(x2, x2);
    const __m512d x8 = _mm512_mul_pd(x4, x4);
    const __m512d poly3 =
        _mm512_fmadd_pd(x4, _mm512_set1_pd(C_6), _mm512_fmadd_pd(x2, _mm512_set1_pd(C_5), _mm512_set1_pd(C_4)));
    const __m512d poly4 =
        _mm512_fmadd_pd(x4, _mm512_fmadd_pd(x2, _mm512_set1_pd(C_3), _mm512_set1_pd(C_2)),
                        _mm512_fmadd_pd(x2, _mm512_set1_pd(C_1), _mm512_set1_pd(C_0)));
    const __m512d t = _mm512_fmadd_pd(x8, poly3, poly4);
    const __m512d hi = _mm512_set1_pd(LOG_2_HI), lo = _mm512_set1_pd(LOG_2_LO);
    const __m512d sp = _mm512_mul_pd(hi, e);
    const __m512d sr = _mm512_fmsub_pd(hi, e, sp);
    const __m512d se = _mm512_fmadd_pd(lo, e, sr);
    const __m512d bx0 = _mm512_add_pd(q0, q0), bx1 = _mm512_add_pd(q1, q1);
    const __m512d s1 = _mm512_add_pd(sp, bx0);
    const __m512d z1 = _mm512_sub_pd(s1, sp);
    const __m512d t1 = _mm512_sub_pd(bx0, z1);
    const __m512d e1 = _mm512_add_pd(t1, _mm512_add_pd(se, bx1));
    const __m512d bb = _mm512_mul_pd(_mm512_mul_pd(x2, q0), t);
    const __m512d s2 = _mm512_add_pd(s1, bb);
    const __m512d z2 = _mm512_sub_pd(s2, s1);
    const __m512d t2 = _mm512_sub_pd(bb, z2);
    const __m512d e2 = _mm512_add_pd(t2, e1);
    __m512d res = _mm512_add_pd(s2, e2);
    const __m512d zero = _mm512_setzero_pd();
    res = _mm512_mask_blend_pd(_mm512_cmp_pd_mask(d, zero, _CMP_NGE_UQ), res, _mm512_set1_pd(NAN));
    res = _mm512_mask_blend_pd(_mm512_cmp_pd_mask(d, _mm512_set1_pd(INFINITY), _CMP_EQ_OQ), res,
                               _mm512_set1_pd(INFINITY));
    res = _mm512_mask_blend_pd(_mm512_cmp_pd_mask(d, zero, _CMP_EQ_OQ), res, _mm512_set1_pd(-INFINITY));
    return res;
}

double ko_lane_log8_fast(double d, int lane) {
    double v[8], o[8];
    for (int j = 0; j < 8; j++) v[j] = j == lane ? d : 1.0;
    _mm512_storeu_pd(o, log8(_mm512_loadu_pd(v)));
    return o[lane];
}

/* _mm512_reduce_add_pd's tree (simd/avx512.rs:104-106): 256-bit halves, then
 * the 128-bit halves, then the pair. */
static inline double radd(__m512d v) {
    double a[8], b[4], c[2];
    _mm512_storeu_pd(a, v);
    for (int i = 0; i < 4; i++) b[i] = a[i] + a[i + 4];
    for (int i = 0; i < 2; i++) c[i] = b[i] + b[i + 2];
    return c[0] + c[1];
}

double ko_jacq_obj_lv_fast(const ko_lanevec* lv, const double x[9], double eps) {
    __m512d zx[9];
    for (int c = 0; c < 9; c++) zx[c] = _mm512_set1_pd(x[c]);
    const __m512d ze = _mm512_set1_pd(eps);
    __m512d zs = _mm512_setzero_pd();
    for (size_t b = 0; b < lv->nblocks; b++) {
        const double* blk = lv->blocks + b * 72;
        __m512d d = ze;
        for (int c = 0; c < 9; c++) d = _mm512_fmadd_pd(zx[c], _mm512_loadu_pd(blk + c * 8), d);
        zs = _mm512_add_pd(zs, log8(d));
    }
    const double bsum = radd(zs);
    double rs = 0.0;
    for (size_t r = 0; r < lv->nrem; r++) {
        double d = eps;
        for (int c = 0; c < 9; c++) d = fma(x[c], lv->rem[r * 9 + c], d);
        rs = rs + ko_lane_log_avx512(d);
    }
    const double n = (double)(lv->nblocks * 8 + lv->nrem);
    return -(bsum + rs) / n;
}

#define TILE 32

/* Hessian pass 1: entries (0,0), (0,1), (0,2), (0,3), (0,4), (0,5), (0,6), (0,7), (0,8), (1,1), (1,2), (1,3), (1,4), (1,5), (1,6) */
static inline void hess_pass1(const __m512d (*buf)[9], size_t nb, __m512d* h) {
    __m512d z0, z1, z2, z3, z4, z5, z6, z7, z8, z9, z10, z11, z12, z13, z14;
    z0 = z1 = z2 = z3 = z4 = z5 = z6 = z7 = z8 = z9 = z10 = z11 = z12 = z13 = z14 = _mm512_setzero_pd();
    for (size_t b = 0; b < nb; b++) {
        const __m512d* c = buf[b];
        z0 = _mm512_fmadd_pd(c[0], c[0], z0);
        z1 = _mm512_fmadd_pd(c[0], c[1], z1);
        z2 = _mm512_fmadd_pd(c[0], c[2], z2);
        z3 = _mm512_fmadd_pd(c[0], c[3], z3);
        z4 = _mm512_fmadd_pd(c[0], c[4], z4);
        z5 = _mm512_fmadd_pd(c[0], c[5], z5);
        z6 = _mm512_fmadd_pd(c[0], c[6], z6);
        z7 = _mm512_fmadd_pd(c[0], c[7], z7);
        z8 = _mm512_fmadd_pd(c[0], c[8], z8);
        z9 = _mm512_fmadd_pd(c[1], c[1], z9);
        z10 = _mm512_fmadd_pd(c[1], c[2], z10);
        z11 = _mm512_fmadd_pd(c[1], c[3], z11);
        z12 = _mm512_fmadd_pd(c[1], c[4], z12);
        z13 = _mm512_fmadd_pd(c[1], c[5], z13);
        z14 = _mm512_fmadd_pd(c[1], c[6], z14);
    }
    h[0] = _mm512_add_pd(h[0], z0);
    h[1] = _mm512_add_pd(h[1], z1);
    h[2] = _mm512_add_pd(h[2], z2);
    h[3] = _mm512_add_pd(h[3], z3);
    h[4] = _mm512_add_pd(h[4], z4);
    h[5] = _mm512_add_pd(h[5], z5);
    h[6] = _mm512_add_pd(h[6], z6);
    h[7] = _mm512_add_pd(h[7], z7);
    h[8] = _mm512_add_pd(h[8], z8);
    h[9] = _mm512_add_pd(h[9], z9);
    h[10] = _mm512_add_pd(h[10], z10);
    h[11] = _mm512_add_pd(h[11], z11);
    h[12] = _mm512_add_pd(h[12], z12);
    h[13] = _mm512_add_pd(h[13], z13);
    h[14] = _mm512_add_pd(h[14], z14);
}

/* Hessian pass 2: entries (1,7), (1,8), (2,2), (2,3), (2,4), (2,5), (2,6), (2,7), (2,8), (3,3), (3,4), (3,5), (3,6), (3,7), (3,8) */
static inline void hess_pass2(const __m512d (*buf)[9], size_t nb, __m512d* h) {
    __m512d z0, z1, z2, z3, z4, z5, z6, z7, z8, z9, z10, z11, z12, z13, z14;
    z0 = z1 = z2 = z3 = z4 = z5 = z6 = z7 = z8 = z9 = z10 = z11 = z12 = z13 = z14 = _mm512_setzero_pd();
    for (size_t b = 0; b < nb; b++) {
        const __m512d* c = buf[b];
        z0 = _mm512_fmadd_pd(c[1], c[7], z0);
        z1 = _mm512_fmadd_pd(c[1], c[8], z1);
        z2 = _mm512_fmadd_pd(c[2], c[2], z2);
        z3 = _mm512_fmadd_pd(c[2], c[3], z3);
        z4 = _mm512_fmadd_pd(c[2], c[4], z4);
        z5 = _mm512_fmadd_pd(c[2], c[5], z5);
        z6 = _mm512_fmadd_pd(c[2], c[6], z6);
        z7 = _mm512_fmadd_pd(c[2], c[7], z7);
        z8 = _mm512_fmadd_pd(c[2], c[8], z8);
        z9 = _mm512_fmadd_pd(c[3], c[3], z9);
        z10 = _mm512_fmadd_pd(c[3], c[4], z10);
        z11 = _mm512_fmadd_pd(c[3], c[5], z11);
        z12 = _mm512_fmadd_pd(c[3], c[6], z12);
        z13 = _mm512_fmadd_pd(c[3], c[7], z13);
        z14 = _mm512_fmadd_pd(c[3], c[8], z14);
    }
    h[15] = _mm512_add_pd(h[15], z0);
    h[16] = _mm512_add_pd(h[16], z1);
    h[17] = _mm512_add_pd(h[17], z2);
    h[18] = _mm512_add_pd(h[18], z3);
    h[19] = _mm512_add_pd(h[19], z4);
    h[20] = _mm512_add_pd(h[20], z5);
    h[21] = _mm512_add_pd(h[21], z6);
    h[22] = _mm512_add_pd(h[22], z7);
    h[23] = _mm512_add_pd(h[23], z8);
    h[24] = _mm512_add_pd(h[24], z9);
    h[25] = _mm512_add_pd(h[25], z10);
    h[26] = _mm512_add_pd(h[26], z11);
    h[27] = _mm512_add_pd(h[27], z12);
    h[28] = _mm512_add_pd(h[28], z13);
    h[29] = _mm512_add_pd(h[29], z14);
}

/* Hessian pass 3: entries (4,4), (4,5), (4,6), (4,7), (4,8), (5,5), (5,6), (5,7), (5,8), (6,6), (6,7), (6,8), (7,7), (7,8), (8,8) */
static inline void hess_pass3(const __m512d (*buf)[9], size_t nb, __m512d* h) {
    __m512d z0, z1, z2, z3, z4, z5, z6, z7, z8, z9, z10, z11, z12, z13, z14;
    z0 = z1 = z2 = z3 = z4 = z5 = z6 = z7 = z8 = z9 = z10 = z11 = z12 = z13 = z14 = _mm512_setzero_pd();
    for (size_t b = 0; b < nb; b++) {
        const __m512d* c = buf[b];
        z0 = _mm512_fmadd_pd(c[4], c[4], z0);
        z1 = _mm512_fmadd_pd(c[4], c[5], z1);
        z2 = _mm512_fmadd_pd(c[4], c[6], z2);
        z3 = _mm512_fmadd_pd(c[4], c[7], z3);
        z4 = _mm512_fmadd_pd(c[4], c[8], z4);
        z5 = _mm512_fmadd_pd(c[5], c[5], z5);
        z6 = _mm512_fmadd_pd(c[5], c[6], z6);
        z7 = _mm512_fmadd_pd(c[5], c[7], z7);
        z8 = _mm512_fmadd_pd(c[5], c[8], z8);
        z9 = _mm512_fmadd_pd(c[6], c[6], z9);
        z10 = _mm512_fmadd_pd(c[6], c[7], z10);
        z11 = _mm512_fmadd_pd(c[6], c[8], z11);
        z12 = _mm512_fmadd_pd(c[7], c[7], z12);
        z13 = _mm512_fmadd_pd(c[7], c[8], z13);
        z14 = _mm512_fmadd_pd(c[8], c[8], z14);
    }
    h[30] = _mm512_add_pd(h[30], z0);
    h[31] = _mm512_add_pd(h[31], z1);
    h[32] = _mm512_add_pd(h[32], z2);
    h[33] = _mm512_add_pd(h[33], z3);
    h[34] = _mm512_add_pd(h[34], z4);
    h[35] = _mm512_add_pd(h[35], z5);
    h[36] = _mm512_add_pd(h[36], z6);
    h[37] = _mm512_add_pd(h[37], z7);
    h[38] = _mm512_add_pd(h[38], z8);
    h[39] = _mm512_add_pd(h[39], z9);
    h[40] = _mm512_add_pd(h[40], z10);
    h[41] = _mm512_add_pd(h[41], z11);
    h[42] = _mm512_add_pd(h[42], z12);
    h[43] = _mm512_add_pd(h[43], z13);
    h[44] = _mm512_add_pd(h[44], z14);
}

/* Entry (a, c) of upper-triangle index id, a <= c. */
static const unsigned char AC[45][2] = {
    {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {0, 7}, {0, 8},
    {1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {1, 6}, {1, 7}, {1, 8},
    {2, 2}, {2, 3}, {2, 4}, {2, 5}, {2, 6}, {2, 7}, {2, 8},
    {3, 3}, {3, 4}, {3, 5}, {3, 6}, {3, 7}, {3, 8},
    {4, 4}, {4, 5}, {4, 6}, {4, 7}, {4, 8},
    {5, 5}, {5, 6}, {5, 7}, {5, 8},
    {6, 6}, {6, 7}, {6, 8},
    {7, 7}, {7, 8},
    {8, 8},
};

void ko_jacq_gh_lv_fast(const ko_lanevec* lv, const double x[9], double eps, double gout[9], double hout[81]) {
    static __thread __m512d buf[TILE][9];
    __m512d zx[9];
    for (int c = 0; c < 9; c++) zx[c] = _mm512_set1_pd(x[c]);
    const __m512d ze = _mm512_set1_pd(eps), one = _mm512_set1_pd(1.0);
    __m512d g[9], h[45];
    for (int c = 0; c < 9; c++) g[c] = _mm512_setzero_pd();
    for (int k = 0; k < 45; k++) h[k] = _mm512_setzero_pd();
    for (size_t t0 = 0; t0 < lv->nblocks || t0 == 0; t0 += TILE) {
        const size_t nb = lv->nblocks - t0 < TILE ? lv->nblocks - t0 : TILE;
        __m512d zg[9];
        for (int c = 0; c < 9; c++) zg[c] = _mm512_setzero_pd();
        for (size_t b = 0; b < nb; b++) {
            const double* blk = lv->blocks + (t0 + b) * 72;
            __m512d c[9];
            for (int k = 0; k < 9; k++) c[k] = _mm512_loadu_pd(blk + k * 8);
            __m512d d = ze;
            for (int k = 0; k < 9; k++) d = _mm512_fmadd_pd(zx[k], c[k], d);
            d = _mm512_div_pd(one, d);
            for (int k = 0; k < 9; k++) zg[k] = _mm512_fmadd_pd(c[k], d, zg[k]);
            for (int k = 0; k < 9; k++) buf[b][k] = _mm512_mul_pd(c[k], d);
        }
        for (int c = 0; c < 9; c++) g[c] = _mm512_add_pd(g[c], zg[c]);
        hess_pass1((const __m512d(*)[9])buf, nb, h);
        hess_pass2((const __m512d(*)[9])buf, nb, h);
        hess_pass3((const __m512d(*)[9])buf, nb, h);
        if (nb < TILE) break;  /* the partial tile (possibly empty) was the last */
    }
    double bg[9], bh[45];
    for (int c = 0; c < 9; c++) bg[c] = radd(g[c]);
    for (int k = 0; k < 45; k++) bh[k] = radd(h[k]);

    /* remainder rows: the scalar lane, one partial tile */
    double rzg[9] = {0}, rbuf[8][9];
    for (size_t r = 0; r < lv->nrem; r++) {
        const double* row = lv->rem + r * 9;
        double d = eps;
        for (int c = 0; c < 9; c++) d = fma(x[c], row[c], d);
        d = 1.0 / d;
        for (int c = 0; c < 9; c++) rzg[c] = fma(row[c], d, rzg[c]);
        for (int c = 0; c < 9; c++) rbuf[r][c] = row[c] * d;
    }
    double rg[9], rh[45];
    for (int c = 0; c < 9; c++) rg[c] = 0.0 + rzg[c];
    for (int id = 0; id < 45; id++) {
        double z = 0.0;
        for (size_t r = 0; r < lv->nrem; r++) z = fma(rbuf[r][AC[id][0]], rbuf[r][AC[id][1]], z);
        rh[id] = 0.0 + z;
    }
    const double n = (double)(lv->nblocks * 8 + lv->nrem);
    for (int c = 0; c < 9; c++) gout[c] = -(bg[c] + rg[c]) / n;
    for (int id = 0; id < 45; id++) {
        const double v = (bh[id] + rh[id]) / n;
        hout[AC[id][0] * 9 + AC[id][1]] = v;
        hout[AC[id][1] * 9 + AC[id][0]] = v;
    }
}

/* mixture_row2 (mixture.rs:96-236) on 8 sites: lx, ly one-hot likelihood
 * vectors (NaN for a missing call), p, q frequencies; a[9] the c_s rows. */
static inline void mixture8(const __m512d lx[3], const __m512d ly[3], __m512d p, __m512d q, __m512d a[9]) {
#define M(x, y) _mm512_mul_pd(x, y)
#define A(x, y) _mm512_add_pd(x, y)
#define F(g, k, acc) acc = _mm512_fmadd_pd(g, k, acc)
    const __m512d p2 = M(p, p), p3 = M(p2, p), q2 = M(q, q), q3 = M(q2, q);
    const __m512d pq = M(p, q), p2q = M(p2, q), pq2 = M(p, q2), p2q2 = M(p2, q2);
    const __m512d p3q = M(p3, q), pq3 = M(p, q3);
    const __m512d twopq = A(pq, pq), twop2q = A(p2q, p2q), twop3q = A(p3q, p3q), twopq2 = A(pq2, pq2),
                  twopq3 = A(pq3, pq3);
    for (int k = 0; k < 9; k++) a[k] = _mm512_setzero_pd();
    __m512d g;
    g = M(lx[0], ly[0]);
    F(g, p, a[0]); F(g, p2, a[1]); F(g, p2, a[2]); F(g, p3, a[3]); F(g, p2, a[4]); F(g, p3, a[5]);
    F(g, p2, a[6]); F(g, p3, a[7]); F(g, M(p2, p2), a[8]);
    g = M(lx[0], ly[1]);
    F(g, pq, a[2]); F(g, twop2q, a[3]); F(g, p2q, a[7]); F(g, twop3q, a[8]);
    g = M(lx[0], ly[2]);
    F(g, pq, a[1]); F(g, pq2, a[3]); F(g, p2q, a[5]); F(g, p2q2, a[8]);
    g = M(lx[1], ly[0]);
    F(g, pq, a[4]); F(g, twop2q, a[5]); F(g, p2q, a[7]); F(g, twop3q, a[8]);
    g = M(lx[1], ly[1]);
    F(g, twopq, a[6]); F(g, pq, a[7]); F(g, M(twopq, twopq), a[8]);
    g = M(lx[1], ly[2]);
    F(g, pq, a[4]); F(g, twopq2, a[5]); F(g, pq2, a[7]); F(g, twopq3, a[8]);
    g = M(lx[2], ly[0]);
    F(g, pq, a[1]); F(g, p2q, a[3]); F(g, pq2, a[5]); F(g, p2q2, a[8]);
    g = M(lx[2], ly[1]);
    F(g, pq, a[2]); F(g, twopq2, a[3]); F(g, pq2, a[7]); F(g, twopq3, a[8]);
    g = M(lx[2], ly[2]);
    F(g, q, a[0]); F(g, q2, a[1]); F(g, q2, a[2]); F(g, q3, a[3]); F(g, q2, a[4]); F(g, q3, a[5]);
    F(g, q2, a[6]); F(g, q3, a[7]); F(g, M(q2, q2), a[8]);
#undef M
#undef A
#undef F
}

/* One-hot likelihood vectors of 8 genotype codes (3 = missing: NaN). */
static inline void onehot8(const uint8_t* codes, __m512d l[3]) {
    const __m128i c8 = _mm_loadl_epi64((const __m128i*)codes);
    const __m512i c = _mm512_cvtepu8_epi64(c8);
    const __mmask8 miss = _mm512_cmpeq_epi64_mask(c, _mm512_set1_epi64(3));
    for (int k = 0; k < 3; k++) {
        const __mmask8 hit = _mm512_cmpeq_epi64_mask(c, _mm512_set1_epi64(k));
        l[k] = _mm512_mask_blend_pd(miss, _mm512_maskz_mov_pd(hit, _mm512_set1_pd(1.0)), _mm512_set1_pd(NAN));
    }
}

/* The pair's c_s rows over V sites as LaneVectorBuilder fills them: each
 * 8-site block through mixture_row2, rows pushed in site order unless they
 * hold a NaN; the last V mod 8 sites one at a time (the scalar lane).
 * Returns the usable rows n_xy; lv->blocks holds room for V / 8 + 1 blocks. */
size_t ko_pair_lanes_codes_fast(const uint8_t* gx, const uint8_t* gy, const double* p, const double* q, size_t V,
                                ko_lanevec* lv) {
    size_t m = 0;
    double* blocks = lv->blocks;
    double row[8][9];
    const size_t V8 = V / 8 * 8;
    for (size_t s = 0; s < V8; s += 8) {
        __m512d lx[3], ly[3], a[9];
        onehot8(gx + s, lx);
        onehot8(gy + s, ly);
        mixture8(lx, ly, _mm512_loadu_pd(p + s), _mm512_loadu_pd(q + s), a);
        __mmask8 ok = 0xff;
        for (int k = 0; k < 9; k++) ok &= _mm512_cmp_pd_mask(a[k], a[k], _CMP_ORD_Q);
        if (ok == 0xff && m % 8 == 0) {  /* a whole block lands on a block boundary */
            for (int k = 0; k < 9; k++) _mm512_storeu_pd(blocks + (m / 8) * 72 + k * 8, a[k]);
            m += 8;
            continue;
        }
        for (int k = 0; k < 9; k++) {
            double t[8];
            _mm512_storeu_pd(t, a[k]);
            for (int j = 0; j < 8; j++) row[j][k] = t[j];
        }
        for (int j = 0; j < 8; j++) {
            if (!((ok >> j) & 1)) continue;
            double* dst = blocks + (m / 8) * 72 + (m % 8);
            for (int k = 0; k < 9; k++) dst[k * 8] = row[j][k];
            m++;
        }
    }
    for (size_t s = V8; s < V; s++) {
        if (gx[s] == 3 || gy[s] == 3) continue;
        /* the scalar lane: mixture_row2 on one site (the same operations) */
        double lx1[3] = {0, 0, 0}, ly1[3] = {0, 0, 0};
        lx1[gx[s]] = 1.0;
        ly1[gy[s]] = 1.0;
        double v[8];
        __m512d lx[3], ly[3], a[9];
        for (int k = 0; k < 3; k++) {
            lx[k] = _mm512_set1_pd(lx1[k]);
            ly[k] = _mm512_set1_pd(ly1[k]);
        }
        mixture8(lx, ly, _mm512_set1_pd(p[s]), _mm512_set1_pd(q[s]), a);
        double* dst = blocks + (m / 8) * 72 + (m % 8);
        for (int k = 0; k < 9; k++) {
            _mm512_storeu_pd(v, a[k]);
            dst[k * 8] = v[0];
        }
        m++;
    }
    /* finish(): the last m mod 8 rows become the scalar remainder */
    lv->nblocks = m / 8;
    lv->nrem = m % 8;
    for (size_t r = 0; r < lv->nrem; r++)
        for (int k = 0; k < 9; k++) lv->rem[r * 9 + k] = blocks[lv->nblocks * 72 + k * 8 + r];
    return m;
}
