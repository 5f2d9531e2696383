/*
 * ko_lanes.c -- the 8-lane per-site loops of the oracle (test infrastructure
 * only; see kestrel_oracle.h).  This file is compiled once per x86 ISA level
 * (AVX-512, AVX2+FMA, baseline) with KO_SUFFIX naming the variant; the
 * arithmetic is identical IEEE FP64 with explicit fma() and no contraction,
 * so every variant returns the same bits.  ko_dispatch.c picks the widest
 * variant the host supports, as simd/simd.rs:15-50 does.
 *
 * Layout follows lanevector.rs:3-78: a pair's compacted rows are stored as
 * blocks of [9][8] doubles (column-major within the block, one double per
 * lane) plus up to 7 scalar remainder rows.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ko_internal.h"

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, KO_SUFFIX)

typedef union {
    double f;
    uint64_t u;
} ko_bits;

/* log/mod.rs:3-14 constants */
#define D1_32 4294967296.0
#define LOG_2_HI 0.693147180559945286226764
#define LOG_2_LO 2.319046813846299558417771e-17
#define C_6 0.1532076988502701353
#define C_5 0.1525629051003428716
#define C_4 0.1818605932937785996
#define C_3 0.2222214519839380009
#define C_2 0.2857142932794299317
#define C_1 0.399999999963525199
#define C_0 0.6666666666667333541

/* log/scalar.rs:17-70, the SLEEF u10 method restated:
 * d = m * 2^e with m in [0.75, 1.5); x = (m-1)/(m+1) as a double-word;
 * log d = e*ln2 + 2x + x^3 * P(x^2). Written branch-free so the 8-lane loops
 * vectorize. */
static inline double lane_log(double d0) {
    const int o = d0 < 2.2250738585072014e-308;
    double d = o ? d0 * (D1_32 * D1_32) : d0;
    ko_bits b;
    b.f = d * (1.0 / 0.75);
    int64_t e = (int64_t)((b.u >> 52) & 0x7ff) - 0x3ff; /* ilogb2k, log/scalar.rs:61-64 */
    ko_bits mb;
    mb.f = d;
    mb.u = (uint64_t)((int64_t)mb.u - (e << 52)); /* ldexp3k, log/scalar.rs:66-70 */
    const double m = mb.f;
    if (o) e -= 64;

    const double u = m - 1.0;
    /* fast_two_sum_ss(1.0, m), log/mod.rs:31-37 */
    const double l0 = 1.0 + m;
    const double lz = l0 - 1.0;
    const double l1 = m - lz;
    /* div_sd(u, l), log/mod.rs:72-85 */
    const double r = 1.0 / l0;
    const double q0 = u * r;
    const double uu = fma(u, r, -q0);
    const double ee = fma(-l0, r, 1.0);
    double q1 = fma(-l1, r, ee);
    q1 = fma(q0, q1, uu);
    const double x0 = q0, x1 = q1;

    const double x2 = x0 * x0;
    const double x4 = x2 * x2;
    const double x8 = x4 * x4;
    /* poly7(x2, x4, x8, C_6..C_0), log/mod.rs:87-100 */
    const double poly3 = fma(x4, C_6, fma(x2, C_5, C_4));
    const double poly4 = fma(x4, fma(x2, C_3, C_2), fma(x2, C_1, C_0));
    const double t = fma(x8, poly3, poly4);

    /* mul_ds((LOG_2_HI, LOG_2_LO), e) */
    const double ef = (double)e;
    const double sp = LOG_2_HI * ef;
    const double sr = fma(LOG_2_HI, ef, -sp);
    const double se = fma(LOG_2_LO, ef, sr);
    /* fast_two_sum_dd(s, (x0 + x0, x1 + x1)) */
    const double bx0 = x0 + x0, bx1 = x1 + x1;
    const double s1 = sp + bx0;
    const double z1 = s1 - sp;
    const double t1 = bx0 - z1;
    const double e1 = t1 + (se + bx1);
    /* fast_two_sum_ds(s, (x2 * x0) * t) */
    const double bb = (x2 * x0) * t;
    const double s2 = s1 + bb;
    const double z2 = s2 - s1;
    const double t2 = bb - z2;
    const double e2 = t2 + e1;
    double res = s2 + e2;

    /* log/scalar.rs:49-57 special cases */
    if (d0 != d0 || d0 < 0.0) res = NAN;
    if (d0 == INFINITY) res = INFINITY;
    if (d0 == 0.0) res = -INFINITY;
    return res;
}

double FN(ko_lane_log_)(double d) { return lane_log(d); }

/* simd/avx512.rs:104-106 -> Rust stdarch _mm512_reduce_add_pd:
 * add the 256-bit halves, then the 128-bit halves, then the pair. */
static inline double radd8(const double v[8]) {
    double a[4], b[2];
    for (int i = 0; i < 4; i++) a[i] = v[i] + v[i + 4];
    for (int i = 0; i < 2; i++) b[i] = a[i] + a[i + 2];
    return b[0] + b[1];
}

/* jacquard/objective.rs:42-72 over the lane blocks, radd, plus the scalar
 * remainder; returns -(b + r) / n (objective.rs:34-39). */
double FN(ko_jacq_obj_lv_)(const ko_lanevec* lv, const double x[9], double eps) {
    double zs[8];
    for (int j = 0; j < 8; j++) zs[j] = 0.0;
    for (size_t b = 0; b < lv->nblocks; b++) {
        const double* blk = lv->blocks + b * 72;
        double d[8];
        for (int j = 0; j < 8; j++) d[j] = eps;
        for (int col = 0; col < 9; col++)
            for (int j = 0; j < 8; j++) d[j] = fma(x[col], blk[col * 8 + j], d[j]);
        for (int j = 0; j < 8; j++) zs[j] = zs[j] + lane_log(d[j]);
    }
    const double bsum = radd8(zs);
    double rs = 0.0;
    for (size_t r = 0; r < lv->nrem; r++) {
        double d = eps;
        for (int col = 0; col < 9; col++) d = fma(x[col], lv->rem[r * 9 + col], d);
        rs = rs + lane_log(d);
    }
    const double n = (double)(lv->nblocks * 8 + lv->nrem);
    return -(bsum + rs) / n;
}

/* One tile of jacquard/grad_hess.rs:89-449 for the 8-lane blocks: per-lane
 * partial sums from zero over <= 32 blocks, then added into the running lane
 * sums g[9][8], h[45][8]. */
static void tile8(const double* blocks, size_t nb, const double x[9], double eps, double g[9][8],
                  double h[45][8], double buf[32][9][8]) {
    double zg[9][8];
    for (int c = 0; c < 9; c++)
        for (int j = 0; j < 8; j++) zg[c][j] = 0.0;
    for (size_t b = 0; b < nb; b++) {
        const double* blk = blocks + b * 72;
        double d[8];
        for (int j = 0; j < 8; j++) d[j] = eps;
        for (int col = 0; col < 9; col++)
            for (int j = 0; j < 8; j++) d[j] = fma(x[col], blk[col * 8 + j], d[j]);
        for (int j = 0; j < 8; j++) d[j] = 1.0 / d[j];
        for (int col = 0; col < 9; col++)
            for (int j = 0; j < 8; j++) zg[col][j] = fma(blk[col * 8 + j], d[j], zg[col][j]);
        for (int col = 0; col < 9; col++)
            for (int j = 0; j < 8; j++) buf[b][col][j] = blk[col * 8 + j] * d[j];
    }
    for (int c = 0; c < 9; c++)
        for (int j = 0; j < 8; j++) g[c][j] = g[c][j] + zg[c][j];
    int id = 0;
    for (int a = 0; a < 9; a++) {
        for (int c = a; c < 9; c++, id++) {
            double z[8];
            for (int j = 0; j < 8; j++) z[j] = 0.0;
            for (size_t b = 0; b < nb; b++)
                for (int j = 0; j < 8; j++) z[j] = fma(buf[b][a][j], buf[b][c][j], z[j]);
            for (int j = 0; j < 8; j++) h[id][j] = h[id][j] + z[j];
        }
    }
}

/* jacquard/grad_hess.rs:12-87: lanes part (tiles of 32 blocks, then the
 * partial tile), radd, scalar remainder as one partial tile; g = -(b+r)/n,
 * H = (b+r)/n filled symmetrically. */
void FN(ko_jacq_gh_lv_)(const ko_lanevec* lv, const double x[9], double eps, double gout[9], double hout[81]) {
    static __thread double buf[32][9][8];
    double g[9][8], h[45][8];
    memset(g, 0, sizeof g);
    memset(h, 0, sizeof h);
    size_t full = lv->nblocks / 32, part = lv->nblocks % 32;
    for (size_t t = 0; t < full; t++) tile8(lv->blocks + t * 32 * 72, 32, x, eps, g, h, buf);
    tile8(lv->blocks + full * 32 * 72, part, x, eps, g, h, buf);
    double bg[9], bh[45];
    for (int c = 0; c < 9; c++) bg[c] = radd8(g[c]);
    for (int k = 0; k < 45; k++) bh[k] = radd8(h[k]);

    /* remainder rows: scalar lane, one (partial) tile */
    double rzg[9] = {0}, rbuf[8][9];
    for (size_t r = 0; r < lv->nrem; r++) {
        const double* row = lv->rem + r * 9;
        double d = eps;
        for (int col = 0; col < 9; col++) d = fma(x[col], row[col], d);
        d = 1.0 / d;
        for (int col = 0; col < 9; col++) rzg[col] = fma(row[col], d, rzg[col]);
        for (int col = 0; col < 9; col++) rbuf[r][col] = row[col] * d;
    }
    double rg[9], rh[45];
    for (int c = 0; c < 9; c++) rg[c] = 0.0 + rzg[c];
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int c = a; c < 9; c++, id++) {
            double z = 0.0;
            for (size_t r = 0; r < lv->nrem; r++) z = fma(rbuf[r][a], rbuf[r][c], z);
            rh[id] = 0.0 + z;
        }

    const double n = (double)(lv->nblocks * 8 + lv->nrem);
    for (int c = 0; c < 9; c++) gout[c] = -(bg[c] + rg[c]) / n;
    id = 0;
    for (int a = 0; a < 9; a++)
        for (int c = a; c < 9; c++, id++) {
            double v = (bh[id] + rh[id]) / n;
            hout[a * 9 + c] = v;
            hout[c * 9 + a] = v;
        }
}

/* allele/objective.rs:85-110 compute_obj_lane2 (+ generic2 33-48) over
 * biallelic likelihood rows stored as blocks [3][8] + remainder rows [3]. */
double FN(ko_allele_obj2_)(const double* blocks, size_t nb, const double* rem, size_t nrem, const double x[2],
                           double eps) {
    const double x00 = x[0] * x[0], x11 = x[1] * x[1], x01 = x[0] * x[1];
    const double tx01 = x01 + x01;
    double zs[8];
    for (int j = 0; j < 8; j++) zs[j] = 0.0;
    for (size_t b = 0; b < nb; b++) {
        const double* blk = blocks + b * 24;
        for (int j = 0; j < 8; j++) {
            double p = eps;
            p = fma(x00, blk[j], p);
            p = fma(tx01, blk[8 + j], p);
            p = fma(x11, blk[16 + j], p);
            zs[j] = lane_log(p) + zs[j];
        }
    }
    const double bs = radd8(zs);
    double rs = 0.0;
    for (size_t r = 0; r < nrem; r++) {
        double p = eps;
        p = fma(x00, rem[r * 3 + 0], p);
        p = fma(tx01, rem[r * 3 + 1], p);
        p = fma(x11, rem[r * 3 + 2], p);
        rs = lane_log(p) + rs;
    }
    const double n = (double)(nb * 8 + nrem);
    return -(bs + rs) / n;
}

/* allele/grad_hess.rs:107-166 compute_grad_hess_lane2 + build_grad_hess 83-105 */
static inline void allele_gh_row(double l00, double l01, double l11, double x0, double x1, double eps, double* zg0,
                                 double* zg1, double* zh0, double* zh1, double* zh2) {
    double lx0 = x0 * l00;
    lx0 = fma(x1, l01, lx0);
    double lx1 = x0 * l01;
    lx1 = fma(x1, l11, lx1);
    double d = eps;
    d = fma(x0, lx0, d);
    d = fma(x1, lx1, d);
    const double inv_d = 1.0 / d;
    const double lxd0 = lx0 * inv_d;
    const double lxd1 = lx1 * inv_d;
    const double lxd00 = lxd0 * lxd0;
    *zh0 = fma(2.0, lxd00, *zh0);
    *zh0 = fma(-l00, inv_d, *zh0);
    const double lxd01 = lxd0 * lxd1;
    *zh1 = fma(2.0, lxd01, *zh1);
    *zh1 = fma(-l01, inv_d, *zh1);
    const double lxd11 = lxd1 * lxd1;
    *zh2 = fma(2.0, lxd11, *zh2);
    *zh2 = fma(-l11, inv_d, *zh2);
    *zg0 = fma(lx0, inv_d, *zg0);
    *zg1 = fma(lx1, inv_d, *zg1);
}

void FN(ko_allele_gh2_)(const double* blocks, size_t nb, const double* rem, size_t nrem, const double x[2],
                        double eps, double g[2], double h[4]) {
    double zg0[8], zg1[8], zh0[8], zh1[8], zh2[8];
    for (int j = 0; j < 8; j++) zg0[j] = zg1[j] = zh0[j] = zh1[j] = zh2[j] = 0.0;
    for (size_t b = 0; b < nb; b++) {
        const double* blk = blocks + b * 24;
        for (int j = 0; j < 8; j++)
            allele_gh_row(blk[j], blk[8 + j], blk[16 + j], x[0], x[1], eps, &zg0[j], &zg1[j], &zh0[j], &zh1[j],
                          &zh2[j]);
    }
    double rg0 = 0, rg1 = 0, rh0 = 0, rh1 = 0, rh2 = 0;
    for (size_t r = 0; r < nrem; r++)
        allele_gh_row(rem[r * 3], rem[r * 3 + 1], rem[r * 3 + 2], x[0], x[1], eps, &rg0, &rg1, &rh0, &rh1, &rh2);
    const double n = (double)(nb * 8 + nrem);
    g[0] = -2.0 * (radd8(zg0) + rg0) / n;
    g[1] = -2.0 * (radd8(zg1) + rg1) / n;
    const double h00 = 2.0 * (radd8(zh0) + rh0) / n;
    const double h01 = 2.0 * (radd8(zh1) + rh1) / n;
    const double h11 = 2.0 * (radd8(zh2) + rh2) / n;
    h[0] = h00;
    h[1] = h01;
    h[2] = h01;
    h[3] = h11;
}

/* quadratic.rs:12-365 compute_gram_matrix: lanes part in tiles of 32 blocks
 * (per-lane partials from zero, added into running lane sums), radd, scalar
 * remainder as one partial tile; divided by n (the number of sites, not rows). */
void FN(ko_gram_lv_)(const ko_lanevec* lv, double n, double out[81]) {
    double h[45][8];
    memset(h, 0, sizeof h);
    size_t full = lv->nblocks / 32, part = lv->nblocks % 32;
    for (size_t t = 0; t <= full; t++) {
        const size_t nb = t < full ? 32 : part;
        const double* blocks = lv->blocks + t * 32 * 72;
        int id = 0;
        for (int a = 0; a < 9; a++)
            for (int c = a; c < 9; c++, id++) {
                double z[8];
                for (int j = 0; j < 8; j++) z[j] = 0.0;
                for (size_t b = 0; b < nb; b++)
                    for (int j = 0; j < 8; j++) z[j] = fma(blocks[b * 72 + a * 8 + j], blocks[b * 72 + c * 8 + j], z[j]);
                for (int j = 0; j < 8; j++) h[id][j] = h[id][j] + z[j];
            }
    }
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int c = a; c < 9; c++, id++) {
            double z = 0.0;
            for (size_t r = 0; r < lv->nrem; r++) z = fma(lv->rem[r * 9 + a], lv->rem[r * 9 + c], z);
            const double v = (radd8(h[id]) + (0.0 + z)) / n;
            out[a * 9 + c] = v;
            out[c * 9 + a] = v;
        }
}
