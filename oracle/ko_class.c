/*
 * ko_class.c -- TEST INFRASTRUCTURE ONLY: part of the oracle (the checker that
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare
 * against; the product never links, loads or calls it).
 *
 * The class order (KO_ORDER_CLASS) of the Jacquard objective, gradient and
 * Hessian sums of one pair (jacquard/objective.rs:8-52, grad_hess.rs:12-449),
 * restated in the operation order of the device's class-sorted sweep, so that
 * its records can be compared bit for bit.
 *
 * What it computes is the reference's sums, f = -(1/n) sum log L_s, g = -(1/n)
 * sum c_s / L_s, H = (1/n) sum c_s c_s^T / L_s^2 with L_s = eps + c_s . x, over
 * the sites where both calls are present; c_s is mixture.rs:96-236 on one-hot
 * likelihoods, i.e. the row m(gx, gy; p, q) of the genotype pair.  Every such
 * row has at most four distinct non-zero values, in positions fixed by the
 * genotype pair (mixture.rs:137-227), so the sums are carried per pattern of
 * positions in that small basis and expanded at the end:
 *   HS (0,0),(2,2): c = (a, a^2, a^2, a^3, a^2, a^3, a^2, a^3, a^4), a = p or q
 *   B  (0,1),(2,1): c_{2,3,7,8} = u (1, 2w, w, 2w^2), u = pq, w = p or q
 *   C  (1,0),(1,2): c_{4,5,7,8} = u (1, 2w, w, 2w^2)
 *   D  (0,2),(2,0): c_{1,3,5,8} = u (1, w1, w2, w1 w2), (w1, w2) = (q, p) or (p, q)
 *   E  (1,1):       c_{6,7,8}   = (2u, u, 4u^2)
 * (the values of mixture_row2 in exact arithmetic; the sums differ from the
 * reference's in rounding only: tests/test_oracle_class.py holds them to the
 * reference order within 1e-13 relative).
 *
 * Order: per segment [s0, s1) of the pair's sites, blocks of 8192 sites from
 * s0, taken in three passes -- HS; B and C; D and E -- each walking every
 * block's lists of its patterns in turn, each list in ascending site order;
 * entry j of a list goes to lane j mod 64 in chunk j / 64, and chunks 4g ..
 * 4g + 3 of a list form its group g.  Each of the 64 lanes keeps its own sums
 * from zero (plain additions for the gradient terms, fma for the Hessian
 * terms) and its own log product; after the first group that ends at least 13
 * chunks after the last log (so at most 16 factors) and at the segment's end
 * each lane adds log(product) to its objective sum (a group whose product is not a positive normal number adds
 * its factors' logs one at a time, as the segmented order does).  The 64
 * lanes' 47 sums are added in a butterfly (o = 32 .. 1: v[i] + v[i ^ o]),
 * expanded into g[9] and h[45], and the segments folded left to right.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kestrel_oracle.h"
#include "ko_internal.h"

#define CL_BLOCK 8192
#define CL_WAVE 64
#define CL_LOG 16
#define CL_NACC 47

enum { PAT_HS = 0, PAT_B = 1, PAT_C = 2, PAT_D = 3, PAT_E = 4, PAT_NONE = 7 };

/* pattern of genotype pair (gx, gy) and its flag (the second genotype pair of
 * the pattern: (2,2), (2,1), (1,2), (2,0)) */
static int cl_pattern(int gx, int gy, int* flag) {
    *flag = 0;
    if (gx > 2 || gy > 2) return PAT_NONE;
    if (gx == 1 && gy == 1) return PAT_E;
    if (gx == 1) {
        *flag = gy == 2;
        return PAT_C;
    }
    if (gy == 1) {
        *flag = gx == 2;
        return PAT_B;
    }
    *flag = gx == 2;
    return gx == gy ? PAT_HS : PAT_D;
}

typedef struct {
    double x0, x1, x2, x3, x4, x5, x8, T1, T3, B37, B8, C57, E67, E8;
} cl_coef;

static void cl_coefs(const double x[9], cl_coef* c) {
    c->x0 = x[0];
    c->x1 = x[1];
    c->x2 = x[2];
    c->x3 = x[3];
    c->x4 = x[4];
    c->x5 = x[5];
    c->x8 = x[8];
    c->T1 = ((x[1] + x[2]) + x[4]) + x[6];
    c->T3 = (x[3] + x[5]) + x[7];
    c->B37 = (x[3] + x[3]) + x[7];
    c->B8 = x[8] + x[8];
    c->C57 = (x[5] + x[5]) + x[7];
    c->E67 = (x[6] + x[6]) + x[7];
    c->E8 = c->B8 + c->B8;
}

typedef struct {
    double acc[CL_NACC]; /* 0: objective; 1..46 the pattern sums */
    double prod;
    double gl[CL_LOG];
    int ng, normal;
} cl_lane;

static void cl_flush(cl_lane* ln) {
    if (ln->normal && ln->prod >= DBL_MIN && ln->prod <= DBL_MAX)
        ln->acc[0] = ln->acc[0] + ko_log(ln->prod);
    else
        for (int q = 0; q < ln->ng; q++) ln->acc[0] = ln->acc[0] + ko_log(ln->gl[q]);
    ln->prod = 1.0;
    ln->ng = 0;
    ln->normal = 1;
}

static void cl_factor(cl_lane* ln, double L) {
    ln->prod = ln->prod * L;
    ln->gl[ln->ng++] = L;
    ln->normal = ln->normal && L >= DBL_MIN && L <= DBL_MAX;
}

/* one site of pattern k in a lane */
static void cl_site(cl_lane* ln, int k, int f, double p, double q, const cl_coef* c, double eps) {
    double* A = ln->acc;
    if (k == PAT_HS) {
        const double a = f ? q : p;
        const double L = fma(a, fma(a, fma(a, fma(a, c->x8, c->T3), c->T1), c->x0), eps);
        cl_factor(ln, L);
        const double r = 1.0 / L;
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
    } else if (k == PAT_B || k == PAT_C) {
        const int o = k == PAT_B ? 12 : 20;
        const double u = p * q, w = f ? q : p;
        const double lin = k == PAT_B ? c->B37 : c->C57, con = k == PAT_B ? c->x2 : c->x4;
        const double L = fma(u, fma(w, fma(w, c->B8, lin), con), eps);
        cl_factor(ln, L);
        const double r = 1.0 / L;
        const double b0 = u * r, b1 = b0 * w, b2 = b1 * w;
        A[o] = A[o] + b0;
        A[o + 1] = A[o + 1] + b1;
        A[o + 2] = A[o + 2] + b2;
        A[o + 3] = fma(b0, b0, A[o + 3]);
        A[o + 4] = fma(b0, b1, A[o + 4]);
        A[o + 5] = fma(b1, b1, A[o + 5]);
        A[o + 6] = fma(b1, b2, A[o + 6]);
        A[o + 7] = fma(b2, b2, A[o + 7]);
    } else if (k == PAT_D) {
        const double u = p * q, w1 = f ? p : q, w2 = f ? q : p;
        const double L = fma(u, fma(w1, c->x3, fma(w2, fma(w1, c->x8, c->x5), c->x1)), eps);
        cl_factor(ln, L);
        const double r = 1.0 / L;
        double d[4];
        d[0] = u * r;
        d[1] = d[0] * w1;
        d[2] = d[0] * w2;
        d[3] = d[1] * w2;
        for (int i = 0; i < 4; i++) A[28 + i] = A[28 + i] + d[i];
        int id = 32;
        for (int i = 0; i < 4; i++)
            for (int j = i; j < 4; j++, id++) A[id] = fma(d[i], d[j], A[id]);
    } else { /* PAT_E */
        const double u = p * q;
        const double L = fma(u, fma(u, c->E8, c->E67), eps);
        cl_factor(ln, L);
        const double r = 1.0 / L;
        const double e0 = u * r, e1 = e0 * u;
        A[42] = A[42] + e0;
        A[43] = A[43] + e1;
        A[44] = fma(e0, e0, A[44]);
        A[45] = fma(e0, e1, A[45]);
        A[46] = fma(e1, e1, A[46]);
    }
}

/* position -> group of the HS pattern, and (factor, exponent) in B, C, E;
 * index in D (-1: not in the pattern) */
static const int HS_GRP[9] = {0, 1, 1, 2, 1, 2, 1, 2, 3};
static const int B_F[9] = {0, 0, 1, 2, 0, 0, 0, 1, 2}, B_E[9] = {0, 0, 0, 1, 0, 0, 0, 1, 2};
static const int C_F[9] = {0, 0, 0, 0, 1, 2, 0, 1, 2}, C_E[9] = {0, 0, 0, 0, 0, 1, 0, 1, 2};
static const int D_I[9] = {-1, 0, -1, 1, -1, 2, -1, -1, 3};
static const int E_F[9] = {0, 0, 0, 0, 0, 0, 2, 1, 4}, E_E[9] = {0, 0, 0, 0, 0, 0, 0, 0, 1};

static int cl_tri4(int i, int j) {
    static const int base[4] = {0, 4, 7, 9};
    return base[i] + (j - i);
}

/* the 47 sums of a segment -> F, g[9], h[45] */
void ko_class_expand(const double* R, double* F, double G[9], double H[45]) {
    const double *gA = R + 1, *SA = R + 5, *gB = R + 12, *SB = R + 15, *gC = R + 20, *SC = R + 23;
    const double *gD = R + 28, *SD = R + 32, *gE = R + 42, *SE = R + 44;
    *F = R[0];
    G[0] = gA[0];
    G[1] = gA[1] + gD[0];
    G[2] = gA[1] + gB[0];
    G[3] = (gA[2] + 2.0 * gB[1]) + gD[1];
    G[4] = gA[1] + gC[0];
    G[5] = (gA[2] + 2.0 * gC[1]) + gD[2];
    G[6] = gA[1] + 2.0 * gE[0];
    G[7] = ((gA[2] + gB[1]) + gC[1]) + gE[0];
    G[8] = (((gA[3] + 2.0 * gB[2]) + 2.0 * gC[2]) + gD[3]) + 4.0 * gE[1];
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int b = a; b < 9; b++, id++) {
            double h = SA[HS_GRP[a] + HS_GRP[b]];
            if (B_F[a] && B_F[b]) h = h + (double)(B_F[a] * B_F[b]) * SB[B_E[a] + B_E[b]];
            if (C_F[a] && C_F[b]) h = h + (double)(C_F[a] * C_F[b]) * SC[C_E[a] + C_E[b]];
            if (D_I[a] >= 0 && D_I[b] >= 0) h = h + SD[cl_tri4(D_I[a], D_I[b])];
            if (E_F[a] && E_F[b]) h = h + (double)(E_F[a] * E_F[b]) * SE[E_E[a] + E_E[b]];
            H[id] = h;
        }
}

/* Sums of one pair in the class order: gx, gy its codes (0..2, 3 missing)
 * and af (p, q) per site; seg has nseg + 1 site bounds. */
void ko_class_sums(const uint8_t* gx, const uint8_t* gy, const double* af, const uint32_t* seg, size_t nseg,
                   const double x[9], double eps, double* fs, double gs[9], double hs[45], size_t* n_out) {
    cl_coef c;
    cl_coefs(x, &c);
    double F = 0.0, G[9] = {0}, H[45] = {0};
    size_t n = 0;
    cl_lane* ln = (cl_lane*)malloc(sizeof(cl_lane) * CL_WAVE);
    cl_lane* tmp = (cl_lane*)malloc(sizeof(cl_lane) * CL_WAVE);
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * CL_BLOCK);
    uint8_t* flag = (uint8_t*)malloc(CL_BLOCK);
    for (size_t k = 0; k < nseg; k++) {
        const uint32_t s0 = seg[k], s1 = seg[k + 1];
        memset(ln, 0, sizeof(cl_lane) * CL_WAVE);
        for (int l = 0; l < CL_WAVE; l++) {
            ln[l].prod = 1.0;
            ln[l].normal = 1;
        }
        uint64_t since = 0;
        static const int PASS_FIRST[3] = {PAT_HS, PAT_B, PAT_D}, PASS_LAST[3] = {PAT_HS, PAT_C, PAT_E};
        for (int pass = 0; pass < 3; pass++)
        for (uint32_t b0 = s0; b0 < s1; b0 += CL_BLOCK) {
            const uint32_t b1 = b0 + CL_BLOCK < s1 ? b0 + CL_BLOCK : s1;
            for (int pat = PASS_FIRST[pass]; pat <= PASS_LAST[pass]; pat++) {
                uint32_t m = 0;
                for (uint32_t s = b0; s < b1; s++) {
                    int f;
                    if (cl_pattern(gx[s], gy[s], &f) == pat) {
                        list[m] = s;
                        flag[m] = (uint8_t)f;
                        m++;
                    }
                }
                n += m;
                for (uint32_t j0 = 0; j0 < m; j0 += CL_WAVE) {
                    for (int l = 0; l < CL_WAVE; l++) {
                        const uint32_t j = j0 + (uint32_t)l;
                        if (j < m) {
                            const uint32_t s = list[j];
                            cl_site(&ln[l], pat, flag[j], af[2 * (size_t)s], af[2 * (size_t)s + 1], &c, eps);
                        }
                    }
                    /* a group is four chunks (the list's last may have fewer); a
                     * group that ends at least CL_LOG - 3 chunks after the
                     * last log ends the lanes' products */
                    const uint32_t ch = j0 / CL_WAVE;
                    if (ch % 4 == 3 || j0 + CL_WAVE >= m) {
                        since += ch % 4 + 1;
                        if (since >= CL_LOG - 3) {
                            for (int l = 0; l < CL_WAVE; l++) cl_flush(&ln[l]);
                            since = 0;
                        }
                    }
                }
            }
        }
        if (since)
            for (int l = 0; l < CL_WAVE; l++) cl_flush(&ln[l]);
        for (int o = CL_WAVE / 2; o > 0; o >>= 1) {
            for (int l = 0; l < CL_WAVE; l++)
                for (int i = 0; i < CL_NACC; i++) tmp[l].acc[i] = ln[l].acc[i] + ln[l ^ o].acc[i];
            for (int l = 0; l < CL_WAVE; l++) memcpy(ln[l].acc, tmp[l].acc, sizeof ln[l].acc);
        }
        double zf, zg[9], zh[45];
        ko_class_expand(ln[0].acc, &zf, zg, zh);
        if (k == 0) {
            F = zf;
            memcpy(G, zg, sizeof G);
            memcpy(H, zh, sizeof H);
        } else {
            F = F + zf;
            for (int i = 0; i < 9; i++) G[i] = G[i] + zg[i];
            for (int i = 0; i < 45; i++) H[i] = H[i] + zh[i];
        }
    }
    free(ln);
    free(tmp);
    free(list);
    free(flag);
    *fs = F;
    memcpy(gs, G, sizeof G);
    memcpy(hs, H, sizeof H);
    *n_out = n;
}
