/*
 * ko_core.c -- oracle (test infrastructure only; see kestrel_oracle.h):
 * identity-state tables, c_s construction, segmented-order reductions,
 * GMW81 / LDL^T, the active-set QP, the condition estimate, the SQP loop,
 * the allele-frequency fit and the threaded pair loops.  Compiled with
 * -ffp-contract=off: every fused multiply-add below is an explicit fma()
 * exactly where the reference writes mul_add / .fma().
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "kestrel_oracle.h"
#include "ko_internal.h"

/* ---------------------------------------------------------------- dispatch */

static const ko_lane_ops OPS_AVX512 = {ko_lane_log_avx512, ko_jacq_obj_lv_avx512, ko_jacq_gh_lv_avx512,
                                       ko_allele_obj2_avx512, ko_allele_gh2_avx512, ko_gram_lv_avx512, "avx512"};
static const ko_lane_ops OPS_AVX2 = {ko_lane_log_avx2, ko_jacq_obj_lv_avx2, ko_jacq_gh_lv_avx2,
                                     ko_allele_obj2_avx2, ko_allele_gh2_avx2, ko_gram_lv_avx2, "avx2"};
static const ko_lane_ops OPS_BASE = {ko_lane_log_base, ko_jacq_obj_lv_base, ko_jacq_gh_lv_base,
                                     ko_allele_obj2_base, ko_allele_gh2_base, ko_gram_lv_base, "base"};

/* simd/simd.rs:15-50: AVX-512 (F/BW/CD/DQ/VL), else AVX2+FMA, else scalar.
 * KO_ISA=base|avx2|avx512 in the environment pins a variant (for tests). */
const ko_lane_ops* ko_ops(void) {
    static const ko_lane_ops* chosen = NULL;
    if (chosen) return chosen;
    const char* pin = getenv("KO_ISA");
    __builtin_cpu_init();
    int has512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                 __builtin_cpu_supports("avx512cd") && __builtin_cpu_supports("avx512dq") &&
                 __builtin_cpu_supports("avx512vl");
    int has2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    const ko_lane_ops* o = has512 ? &OPS_AVX512 : has2 ? &OPS_AVX2 : &OPS_BASE;
    if (pin) {
        if (!strcmp(pin, "base")) o = &OPS_BASE;
        else if (!strcmp(pin, "avx2") && has2) o = &OPS_AVX2;
        else if (!strcmp(pin, "avx512") && has512) o = &OPS_AVX512;
    }
    chosen = o;
    return o;
}

const char* ko_isa_name(void) { return ko_ops()->name; }

int ko_has_avx512(void) { return ko_ops() == &OPS_AVX512; }

/* ------------------------------------------------------------------ tuning */

/* sqp.rs:31-46 */
void ko_tune_new(ko_tune* t, uint64_t sqp_max_iter, uint64_t qp_max_iter) {
    t->sqp_max_iter = sqp_max_iter;
    t->sqp_conv_tol = 1e-8;
    t->qp_max_iter = qp_max_iter;
    t->qp_conv_tol = 1e-10;
    t->qp_zero_search_tol = 1e-12;
    t->bls_max_iter = 10;
    t->bls_sufficient_decrease = 1e-4;
    t->bls_step_size_reduce = 0.9;
    t->epsilon = 1e-8;
}

/* --------------------------------------------------------------------- log */

double ko_log(double d) { return ko_ops()->log(d); }

/* log/avx512.rs:15-79: getexp / getmant(P75_1P5) form of the same method;
 * kept to check that both reference variants agree on positive normals. */
double ko_log_getmant(double d) {
    if (d != d || d < 0.0) return NAN;
    if (d == 0.0) return -INFINITY;
    if (d == INFINITY) return INFINITY;
    int ex;
    double mant = frexp(d, &ex); /* d = mant * 2^ex, mant in [0.5, 1) */
    double m = mant * 2.0;       /* [1, 2) */
    double e = (double)(ex - 1);
    if (m >= 1.5) m = m * 0.5; /* P75_1P5 */
    if (m < 1.0) e = e + 1.0;
    const double u = m - 1.0;
    const double l0 = 1.0 + m, lz = l0 - 1.0, l1 = m - lz;
    const double r = 1.0 / l0;
    const double q0 = u * r;
    const double uu = fma(u, r, -q0);
    const double ee = fma(-l0, r, 1.0);
    double q1 = fma(-l1, r, ee);
    q1 = fma(q0, q1, uu);
    const double x2 = q0 * q0, x4 = x2 * x2, x8 = x4 * x4;
    const double poly3 = fma(x4, 0.1532076988502701353, fma(x2, 0.1525629051003428716, 0.1818605932937785996));
    const double poly4 = fma(x4, fma(x2, 0.2222214519839380009, 0.2857142932794299317),
                             fma(x2, 0.399999999963525199, 0.6666666666667333541));
    const double t = fma(x8, poly3, poly4);
    const double HI = 0.693147180559945286226764, LO = 2.319046813846299558417771e-17;
    const double sp = HI * e, sr = fma(HI, e, -sp), se = fma(LO, e, sr);
    const double bx0 = q0 + q0, bx1 = q1 + q1;
    const double s1 = sp + bx0, z1 = s1 - sp, t1 = bx0 - z1, e1 = t1 + (se + bx1);
    const double bb = (x2 * q0) * t;
    const double s2 = s1 + bb, z2 = s2 - s1, t2 = bb - z2, e2 = t2 + e1;
    return s2 + e2;
}

/* ----------------------------------------------------- identity states */

/* iis.rs:26-104: six equality bits -> mode 1..15 */
int ko_iis_mode(int i, int j, int k, int l) {
    int key = ((i == j) << 5) | ((i == k) << 4) | ((i == l) << 3) | ((j == k) << 2) | ((j == l) << 1) | (k == l);
    switch (key) {
    case 0x3f: return 1;  /* 111111 */
    case 0x21: return 2;  /* 100001 */
    case 0x34: return 3;  /* 110100 */
    case 0x2a: return 4;  /* 101010 */
    case 0x20: return 5;  /* 100000 */
    case 0x19: return 6;  /* 011001 */
    case 0x07: return 7;  /* 000111 */
    case 0x01: return 8;  /* 000001 */
    case 0x12: return 9;  /* 010010 */
    case 0x0c: return 10; /* 001100 */
    case 0x10: return 11; /* 010000 */
    case 0x08: return 12; /* 001000 */
    case 0x04: return 13; /* 000100 */
    case 0x02: return 14; /* 000010 */
    case 0x00: return 15; /* 000000 */
    default: return 0;    /* not an equivalence relation */
    }
}

/* iis.rs:106-277 accumulate_iis1..15 */
static void accumulate_iis(double a[9], double g, double pi, double pj, double pk, double pl, int mode) {
    switch (mode) {
    case 1: {
        double pi2 = pi * pi, pi3 = pi2 * pi;
        a[0] = fma(g, pi, a[0]);
        a[1] = fma(g, pi2, a[1]);
        a[2] = fma(g, pi2, a[2]);
        a[3] = fma(g, pi3, a[3]);
        a[4] = fma(g, pi2, a[4]);
        a[5] = fma(g, pi3, a[5]);
        a[6] = fma(g, pi2, a[6]);
        a[7] = fma(g, pi3, a[7]);
        a[8] = fma(g, pi2 * pi2, a[8]);
    } break;
    case 2: {
        double pipk = pi * pk;
        a[1] = fma(g, pipk, a[1]);
        a[3] = fma(g, pipk * pk, a[3]);
        a[5] = fma(g, pi * pipk, a[5]);
        a[8] = fma(g, pipk * pipk, a[8]);
    } break;
    case 3: {
        double pipl = pi * pl, pi2pl = pi * pipl, two = pi2pl + pi2pl, two3 = two * pi;
        a[2] = fma(g, pipl, a[2]);
        a[3] = fma(g, two, a[3]);
        a[7] = fma(g, pi2pl, a[7]);
        a[8] = fma(g, two3, a[8]);
    } break;
    case 4: {
        double pipk = pi * pk, pi2pk = pi * pipk, two = pi2pk + pi2pk, two3 = two * pi;
        a[2] = fma(g, pipk, a[2]);
        a[3] = fma(g, two, a[3]);
        a[7] = fma(g, pi2pk, a[7]);
        a[8] = fma(g, two3, a[8]);
    } break;
    case 5: {
        double v = (pi * pk) * pl, two = v + v;
        a[3] = fma(g, two, a[3]);
        a[8] = fma(g, two * pi, a[8]);
    } break;
    case 6: {
        double pipj = pi * pj, pi2pj = pi * pipj, two = pi2pj + pi2pj;
        a[4] = fma(g, pipj, a[4]);
        a[5] = fma(g, two, a[5]);
        a[7] = fma(g, pi2pj, a[7]);
        a[8] = fma(g, two * pi, a[8]);
    } break;
    case 7: {
        double pipj = pi * pj, pj2pi = pj * pipj, two = pj2pi + pj2pi;
        a[4] = fma(g, pipj, a[4]);
        a[5] = fma(g, two, a[5]);
        a[7] = fma(g, pj2pi, a[7]);
        a[8] = fma(g, two * pj, a[8]);
    } break;
    case 8: {
        double v = (pi * pj) * pk, two = v + v;
        a[5] = fma(g, two, a[5]);
        a[8] = fma(g, two * pk, a[8]);
    } break;
    case 9:
    case 10: {
        double pipj = pi * pj, two = pipj + pipj, four = two * two;
        a[6] = fma(g, two, a[6]);
        a[7] = fma(g, pipj * (pi + pj), a[7]);
        a[8] = fma(g, four, a[8]);
    } break;
    case 11: {
        double v = (pi * pj) * pl;
        a[7] = fma(g, v, a[7]);
        a[8] = fma(g, (4.0 * v) * pi, a[8]);
    } break;
    case 12: {
        double v = (pi * pj) * pk;
        a[7] = fma(g, v, a[7]);
        a[8] = fma(g, (4.0 * v) * pi, a[8]);
    } break;
    case 13: {
        double v = (pi * pj) * pl;
        a[7] = fma(g, v, a[7]);
        a[8] = fma(g, (4.0 * v) * pj, a[8]);
    } break;
    case 14: {
        double v = (pi * pj) * pk;
        a[7] = fma(g, v, a[7]);
        a[8] = fma(g, (4.0 * v) * pj, a[8]);
    } break;
    case 15: {
        double v = (pi * pj) * (pk * pl);
        a[8] = fma(g, 4.0 * v, a[8]);
    } break;
    default: break;
    }
}

/* mixture.rs:8-60 over conditional.rs IIS_2 (genotypes in VCF triangular
 * order (0,0),(0,1),(1,1)). */
void ko_mixture_generic2(const double lx[3], const double ly[3], const double af[2], double out[9]) {
    static const int GI[3][2] = {{0, 0}, {0, 1}, {1, 1}};
    double a[9] = {0};
    for (int gx = 0; gx < 3; gx++)
        for (int gy = 0; gy < 3; gy++) {
            double g = lx[gx] * ly[gy];
            int i = GI[gx][0], j = GI[gx][1], k = GI[gy][0], l = GI[gy][1];
            accumulate_iis(a, g, af[i], af[j], af[k], af[l], ko_iis_mode(i, j, k, l));
        }
    memcpy(out, a, sizeof a);
}

/* mixture.rs:8-60 calculate_mixture_component_matrix over IIS_A
 * (conditional.rs:10-50): genotypes in VCF order, index j(j+1)/2 + i for
 * alleles i <= j; for every (gx, gy) in that order g = lx[gx] ly[gy] and
 * accumulate_iis with the four alleles' frequencies. */
static void genotype_alleles(int A, int gi[10], int gj[10]) {
    int g = 0;
    for (int j = 0; j < A; j++)
        for (int i = 0; i <= j; i++, g++) {
            gi[g] = i;
            gj[g] = j;
        }
}

void ko_mixture_generic(int A, const double* lx, const double* ly, const double* af, double out[9]) {
    int gi[10], gj[10];
    genotype_alleles(A, gi, gj);
    const int G = A * (A + 1) / 2;
    double a[9] = {0};
    for (int gx = 0; gx < G; gx++)
        for (int gy = 0; gy < G; gy++) {
            const double g = lx[gx] * ly[gy];
            const int i = gi[gx], j = gj[gx], k = gi[gy], l = gj[gy];
            accumulate_iis(a, g, af[i], af[j], af[k], af[l], ko_iis_mode(i, j, k, l));
        }
    memcpy(out, a, sizeof a);
}

/* iis.rs:279-462 */
void ko_conditional_probabilities(double pi, double pj, double pk, double pl, int mode, double o[9]) {
    const double pi2 = pi * pi, pi3 = pi2 * pi, pj2 = pj * pj, pk2 = pk * pk;
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
    case 10: o[6] = (2.0 * pi) * pj; o[7] = (pi * pj) * (pi + pj); o[8] = (4.0 * pi2) * pj2; break;
    case 11: o[7] = (pi * pj) * pl; o[8] = ((4.0 * pi2) * pj) * pl; break;
    case 12: o[7] = (pi * pj) * pk; o[8] = ((4.0 * pi2) * pj) * pk; break;
    case 13: o[7] = (pi * pj) * pl; o[8] = ((4.0 * pi) * pj2) * pl; break;
    case 14: o[7] = (pi * pj) * pk; o[8] = ((4.0 * pi) * pj2) * pk; break;
    case 15: o[8] = (((4.0 * pi) * pj) * pk) * pl; break;
    default: break;
    }
}

/* mixture.rs:96-236 mixture_row2, in its exact expression order. */
void ko_mixture_row2(const double lx[3], const double ly[3], const double af[2], double out[9]) {
    double a[9] = {0};
    const double p = af[0], q = af[1];
    const double p2 = p * p, p3 = p2 * p, q2 = q * q, q3 = q2 * q;
    const double pq = p * q, p2q = p2 * q, pq2 = p * q2, p2q2 = p2 * q2;
    const double p3q = p3 * q, pq3 = p * q3;
    const double twopq = pq + pq, twop2q = p2q + p2q, twop3q = p3q + p3q, twopq2 = pq2 + pq2, twopq3 = pq3 + pq3;
    double g;
    g = lx[0] * ly[0]; /* (0,0)(0,0) M1 */
    a[0] = fma(g, p, a[0]); a[1] = fma(g, p2, a[1]); a[2] = fma(g, p2, a[2]); a[3] = fma(g, p3, a[3]);
    a[4] = fma(g, p2, a[4]); a[5] = fma(g, p3, a[5]); a[6] = fma(g, p2, a[6]); a[7] = fma(g, p3, a[7]);
    a[8] = fma(g, p2 * p2, a[8]);
    g = lx[0] * ly[1]; /* (0,0)(0,1) M3 */
    a[2] = fma(g, pq, a[2]); a[3] = fma(g, twop2q, a[3]); a[7] = fma(g, p2q, a[7]); a[8] = fma(g, twop3q, a[8]);
    g = lx[0] * ly[2]; /* (0,0)(1,1) M2 */
    a[1] = fma(g, pq, a[1]); a[3] = fma(g, pq2, a[3]); a[5] = fma(g, p2q, a[5]); a[8] = fma(g, p2q2, a[8]);
    g = lx[1] * ly[0]; /* (0,1)(0,0) M6 */
    a[4] = fma(g, pq, a[4]); a[5] = fma(g, twop2q, a[5]); a[7] = fma(g, p2q, a[7]); a[8] = fma(g, twop3q, a[8]);
    g = lx[1] * ly[1]; /* (0,1)(0,1) M9 */
    a[6] = fma(g, twopq, a[6]); a[7] = fma(g, pq, a[7]); a[8] = fma(g, twopq * twopq, a[8]);
    g = lx[1] * ly[2]; /* (0,1)(1,1) M7 */
    a[4] = fma(g, pq, a[4]); a[5] = fma(g, twopq2, a[5]); a[7] = fma(g, pq2, a[7]); a[8] = fma(g, twopq3, a[8]);
    g = lx[2] * ly[0]; /* (1,1)(0,0) M2 */
    a[1] = fma(g, pq, a[1]); a[3] = fma(g, p2q, a[3]); a[5] = fma(g, pq2, a[5]); a[8] = fma(g, p2q2, a[8]);
    g = lx[2] * ly[1]; /* (1,1)(0,1) M4 */
    a[2] = fma(g, pq, a[2]); a[3] = fma(g, twopq2, a[3]); a[7] = fma(g, pq2, a[7]); a[8] = fma(g, twopq3, a[8]);
    g = lx[2] * ly[2]; /* (1,1)(1,1) M1 */
    a[0] = fma(g, q, a[0]); a[1] = fma(g, q2, a[1]); a[2] = fma(g, q2, a[2]); a[3] = fma(g, q3, a[3]);
    a[4] = fma(g, q2, a[4]); a[5] = fma(g, q3, a[5]); a[6] = fma(g, q2, a[6]); a[7] = fma(g, q3, a[7]);
    a[8] = fma(g, q2 * q2, a[8]);
    memcpy(out, a, sizeof a);
}

/* ----------------------------------------------------------- reductions */

static void rows_to_lanevec(const double* rows, size_t n, ko_lanevec* lv) {
    lv->nblocks = n / 8;
    lv->nrem = n % 8;
    lv->blocks = (double*)malloc((lv->nblocks ? lv->nblocks : 1) * 72 * sizeof(double));
    for (size_t b = 0; b < lv->nblocks; b++)
        for (int j = 0; j < 8; j++)
            for (int c = 0; c < 9; c++) lv->blocks[b * 72 + c * 8 + j] = rows[(b * 8 + j) * 9 + c];
    for (size_t r = 0; r < lv->nrem; r++)
        for (int c = 0; c < 9; c++) lv->rem[r * 9 + c] = rows[(lv->nblocks * 8 + r) * 9 + c];
}

double ko_jacq_obj_ref(const double* rows, size_t n, const double x[9], double eps) {
    ko_lanevec lv;
    rows_to_lanevec(rows, n, &lv);
    double f = ko_ops()->jacq_obj(&lv, x, eps);
    free(lv.blocks);
    return f;
}

void ko_jacq_gh_ref(const double* rows, size_t n, const double x[9], double eps, double g[9], double h[81]) {
    ko_lanevec lv;
    rows_to_lanevec(rows, n, &lv);
    ko_ops()->jacq_gh(&lv, x, eps, g, h);
    free(lv.blocks);
}

/* Segmented sequential order (the device's order): per-site operations
 * identical to jacquard/objective.rs:57-68 and grad_hess.rs:107-136, 149-447;
 * each segment sums its usable sites in order from zero and the segment
 * partials are folded left to right.  The objective sum of log L_s is taken
 * as a sum over aligned 16-site groups of log(prod L_s) (usable sites only,
 * product from 1.0 in site order): the same quantity with no more rounding
 * error than 16 separate logs, and one log per group.  A group holding an L
 * that is not a positive normal number, or whose product is not one, adds
 * its sites' logs one at a time in site order instead, so zero, negative,
 * subnormal and infinite L keep the meaning of log/scalar.rs:49-57. */
#define KO_LOG_GROUP 16
static void seg_sums(const double* rows, const uint8_t* valid, const uint32_t* seg, size_t nseg, const double x[9],
                     double eps, int want_obj, int want_gh, double* fs, double gs[9], double hs[45], size_t* n_out) {
    double F = 0.0, G[9] = {0}, H[45] = {0};
    size_t n = 0;
    for (size_t k = 0; k < nseg; k++) {
        double zf = 0.0, zg[9] = {0}, zh[45] = {0};
        double prod = 1.0, gl[KO_LOG_GROUP];
        int ng = 0, normal = 1;
        for (uint32_t s = seg[k]; s < seg[k + 1]; s++) {
            if (valid[s]) {
            n++;
            const double* c = rows + (size_t)s * 9;
            double d = eps;
            for (int col = 0; col < 9; col++) d = fma(x[col], c[col], d);
            if (want_obj) {
                prod = prod * d;
                gl[ng++] = d;
                normal = normal && d >= DBL_MIN && d <= DBL_MAX;
            }
            if (want_gh) {
                const double r = 1.0 / d;
                double cs[9];
                for (int col = 0; col < 9; col++) zg[col] = fma(c[col], r, zg[col]);
                for (int col = 0; col < 9; col++) cs[col] = c[col] * r;
                int id = 0;
                for (int a = 0; a < 9; a++)
                    for (int b = a; b < 9; b++, id++) zh[id] = fma(cs[a], cs[b], zh[id]);
            }
            }
            if ((s + 1) % KO_LOG_GROUP == 0 || s + 1 == seg[k + 1]) {
                if (want_obj) {
                    if (normal && prod >= DBL_MIN && prod <= DBL_MAX)
                        zf = zf + ko_log(prod);
                    else
                        for (int q = 0; q < ng; q++) zf = zf + ko_log(gl[q]);
                }
                prod = 1.0;
                ng = 0;
                normal = 1;
            }
        }
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
    *fs = F;
    memcpy(gs, G, sizeof G);
    memcpy(hs, H, sizeof H);
    *n_out = n;
}

double ko_jacq_obj_seg(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                       const double x[9], double eps) {
    (void)V;
    double F, G[9], H[45];
    size_t n;
    seg_sums(rows, valid, seg, nseg, x, eps, 1, 0, &F, G, H, &n);
    return -F / (double)n;
}

void ko_jacq_gh_seg(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                    const double x[9], double eps, double g[9], double h[81]) {
    (void)V;
    double F, G[9], H[45];
    size_t n;
    seg_sums(rows, valid, seg, nseg, x, eps, 0, 1, &F, G, H, &n);
    const double nn = (double)n;
    for (int i = 0; i < 9; i++) g[i] = -G[i] / nn;
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int b = a; b < 9; b++, id++) {
            double v = H[id] / nn;
            h[a * 9 + b] = v;
            h[b * 9 + a] = v;
        }
}

/* Wavefront order (the device's sweep_lanes_fp64, fits of option
 * "pair_lanes" 64): per segment, "lane" l of 64 takes the sites s0 + l + 64 t
 * (t = 0, 1, ...) and sums its usable sites in that order from zero with the
 * per-site operations of seg_sums; its log product is logged after every 16
 * of its steps and after its last (a step whose site is past the segment or
 * not usable multiplies by nothing, so a group may log 1.0), with seg_sums'
 * per-site fallback where a group is not a positive normal product.  The 64
 * lanes' sums are then added in a butterfly: for o = 32, 16, ..., 1 every
 * lane i takes v[i] + v[i ^ o]; lane 0's sums are the segment's, folded left
 * to right as in seg_sums. */
#define KO_WAVE 64
static void lanes_sums(const double* rows, const uint8_t* valid, const uint32_t* seg, size_t nseg,
                       const double x[9], double eps, int want_obj, int want_gh, double* fs, double gs[9],
                       double hs[45], size_t* n_out) {
    double F = 0.0, G[9] = {0}, H[45] = {0};
    size_t n = 0;
    double (*lf)[1] = malloc(sizeof(double[KO_WAVE][1]));
    double (*lg)[9] = malloc(sizeof(double[KO_WAVE][9]));
    double (*lh)[45] = malloc(sizeof(double[KO_WAVE][45]));
    double (*tf)[1] = malloc(sizeof(double[KO_WAVE][1]));
    double (*tg)[9] = malloc(sizeof(double[KO_WAVE][9]));
    double (*th)[45] = malloc(sizeof(double[KO_WAVE][45]));
    for (size_t k = 0; k < nseg; k++) {
        const uint32_t s0 = seg[k], s1 = seg[k + 1];
        const uint32_t steps = (s1 - s0 + KO_WAVE - 1) / KO_WAVE;
        for (int l = 0; l < KO_WAVE; l++) {
            double zf = 0.0, zg[9] = {0}, zh[45] = {0};
            double prod = 1.0, gl[KO_LOG_GROUP];
            int ng = 0, normal = 1;
            for (uint32_t t = 0; t < steps; t++) {
                const uint32_t s = s0 + KO_WAVE * t + (uint32_t)l;
                if (s < s1 && valid[s]) {
                    n++;
                    const double* c = rows + (size_t)s * 9;
                    double d = eps;
                    for (int col = 0; col < 9; col++) d = fma(x[col], c[col], d);
                    if (want_obj) {
                        prod = prod * d;
                        gl[ng++] = d;
                        normal = normal && d >= DBL_MIN && d <= DBL_MAX;
                    }
                    if (want_gh) {
                        const double r = 1.0 / d;
                        double cs[9];
                        for (int col = 0; col < 9; col++) zg[col] = fma(c[col], r, zg[col]);
                        for (int col = 0; col < 9; col++) cs[col] = c[col] * r;
                        int id = 0;
                        for (int a = 0; a < 9; a++)
                            for (int b = a; b < 9; b++, id++) zh[id] = fma(cs[a], cs[b], zh[id]);
                    }
                }
                if (t % KO_LOG_GROUP == KO_LOG_GROUP - 1 || t + 1 == steps) {
                    if (want_obj) {
                        if (normal && prod >= DBL_MIN && prod <= DBL_MAX)
                            zf = zf + ko_log(prod);
                        else
                            for (int q = 0; q < ng; q++) zf = zf + ko_log(gl[q]);
                    }
                    prod = 1.0;
                    ng = 0;
                    normal = 1;
                }
            }
            lf[l][0] = zf;
            memcpy(lg[l], zg, sizeof zg);
            memcpy(lh[l], zh, sizeof zh);
        }
        for (int o = KO_WAVE / 2; o > 0; o >>= 1) {
            for (int l = 0; l < KO_WAVE; l++) {
                tf[l][0] = lf[l][0] + lf[l ^ o][0];
                for (int i = 0; i < 9; i++) tg[l][i] = lg[l][i] + lg[l ^ o][i];
                for (int i = 0; i < 45; i++) th[l][i] = lh[l][i] + lh[l ^ o][i];
            }
            memcpy(lf, tf, sizeof(double[KO_WAVE][1]));
            memcpy(lg, tg, sizeof(double[KO_WAVE][9]));
            memcpy(lh, th, sizeof(double[KO_WAVE][45]));
        }
        if (k == 0) {
            F = lf[0][0];
            memcpy(G, lg[0], sizeof G);
            memcpy(H, lh[0], sizeof H);
        } else {
            F = F + lf[0][0];
            for (int i = 0; i < 9; i++) G[i] = G[i] + lg[0][i];
            for (int i = 0; i < 45; i++) H[i] = H[i] + lh[0][i];
        }
    }
    free(lf);
    free(lg);
    free(lh);
    free(tf);
    free(tg);
    free(th);
    *fs = F;
    memcpy(gs, G, sizeof G);
    memcpy(hs, H, sizeof H);
    *n_out = n;
}

static void gh_from_sums(const double G[9], const double H[45], size_t n, double g[9], double h[81]) {
    const double nn = (double)n;
    for (int i = 0; i < 9; i++) g[i] = -G[i] / nn;
    int id = 0;
    for (int a = 0; a < 9; a++)
        for (int b = a; b < 9; b++, id++) {
            double v = H[id] / nn;
            h[a * 9 + b] = v;
            h[b * 9 + a] = v;
        }
}

double ko_jacq_obj_lanes(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                         const double x[9], double eps) {
    (void)V;
    double F, G[9], H[45];
    size_t n;
    lanes_sums(rows, valid, seg, nseg, x, eps, 1, 0, &F, G, H, &n);
    return -F / (double)n;
}

void ko_jacq_gh_lanes(const double* rows, const uint8_t* valid, size_t V, const uint32_t* seg, size_t nseg,
                      const double x[9], double eps, double g[9], double h[81]) {
    (void)V;
    double F, G[9], H[45];
    size_t n;
    lanes_sums(rows, valid, seg, nseg, x, eps, 0, 1, &F, G, H, &n);
    gh_from_sums(G, H, n, g, h);
}

/* ------------------------------------------------------ linear algebra */

#define MX 9

/* algebra.rs:4-11 dot (fma chain from zero) */
static double dotn(int n, const double* x, const double* y) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fma(x[i], y[i], s);
    return s;
}

/* algebra.rs:81-84 sum_to_one */
static void sum_to_one(int N, const double* x, double* out) {
    double s = 0.0;
    for (int i = 0; i < N; i++) s += x[i];
    for (int i = 0; i < N; i++) out[i] = x[i] / s;
}

/* Rust f64::total_cmp ordering key */
static int64_t tkey(double v) {
    int64_t k;
    memcpy(&k, &v, 8);
    k ^= (int64_t)(((uint64_t)(k >> 63)) >> 1);
    return k;
}

/* cholesky.rs:6-65 modify_gmw_n (a, l are MX x MX row-major) */
static void gmw(int n, const double* a, double* l, double* d) {
    double gamma = 0.0;
    for (int i = 0; i < n; i++) gamma = fmax(gamma, fabs(a[i * MX + i]));
    double xi = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) xi = fmax(xi, fabs(a[i * MX + j]));
    const double delta = DBL_EPSILON * fmax(1.0, gamma + xi);
    const double t = sqrt((double)n * (double)n - 1.0);
    const double ratio = n <= 1 ? 0.0 : xi / t;
    const double beta2 = fmax(fmax(gamma, ratio), DBL_EPSILON);
    double p[MX] = {0};
    for (int j = 0; j < n; j++) {
        double sjj = 0.0;
        for (int s = 0; s < j; s++) {
            p[s] = d[s] * l[j * MX + s];
            sjj = fma(p[s], l[j * MX + s], sjj);
        }
        const double cjj = a[j * MX + j] - sjj;
        l[j * MX + j] = 1.0;
        double theta = 0.0;
        for (int i = j + 1; i < n; i++) {
            double sij = 0.0;
            for (int s = 0; s < j; s++) sij = fma(p[s], l[i * MX + s], sij);
            const double cij = a[i * MX + j] - sij;
            theta = fmax(theta, fabs(cij));
            l[i * MX + j] = cij;
        }
        const double dj = fmax(fmax(fabs(cjj), theta * theta / beta2), delta);
        for (int i = j + 1; i < n; i++) l[i * MX + j] /= dj;
        d[j] = dj;
    }
}

/* cholesky.rs:72-102 unmodified_n */
static void unmod(int n, const double* a, double* l, double* d) {
    double p[MX] = {0};
    for (int j = 0; j < n; j++) {
        l[j * MX + j] = 1.0;
        double sjj = 0.0;
        for (int s = 0; s < j; s++) {
            p[s] = d[s] * l[j * MX + s];
            sjj = fma(p[s], l[j * MX + s], sjj);
        }
        const double dj = a[j * MX + j] - sjj;
        for (int i = j + 1; i < n; i++) {
            double sij = 0.0;
            for (int s = 0; s < j; s++) sij = fma(p[s], l[i * MX + s], sij);
            const double cij = a[i * MX + j] - sij;
            l[i * MX + j] = cij / dj;
        }
        d[j] = dj;
    }
}

/* cholesky.rs:131-140 */
static void backsub(int n, const double* l, double* b) {
    for (int i = n - 1; i >= 0; i--) {
        double s = 0.0;
        for (int j = i + 1; j < n; j++) s = fma(b[j], l[j * MX + i], s);
        b[i] -= s;
    }
}

/* cholesky.rs:105-128 */
static void ldl_solve(int n, const double* l, const double* d, double* b) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < i; j++) s = fma(b[j], l[i * MX + j], s);
        b[i] -= s;
    }
    for (int i = 0; i < n; i++) b[i] /= d[i];
    backsub(n, l, b);
}

static void to_mx(int n, const double* a, double* o) {
    memset(o, 0, sizeof(double) * MX * MX);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) o[i * MX + j] = a[i * n + j];
}

void ko_modify_gmw(int n, const double* a, double* l, double* d) {
    double A[MX * MX], L[MX * MX] = {0}, D[MX] = {0};
    to_mx(n, a, A);
    gmw(n, A, L, D);
    for (int i = 0; i < n; i++) {
        d[i] = D[i];
        for (int j = 0; j < n; j++) l[i * n + j] = L[i * MX + j];
    }
}

void ko_unmodified(int n, const double* a, double* l, double* d) {
    double A[MX * MX], L[MX * MX] = {0}, D[MX] = {0};
    to_mx(n, a, A);
    unmod(n, A, L, D);
    for (int i = 0; i < n; i++) {
        d[i] = D[i];
        for (int j = 0; j < n; j++) l[i * n + j] = L[i * MX + j];
    }
}

void ko_solve_ldl(int n, const double* l, const double* d, double* b) {
    double L[MX * MX];
    to_mx(n, l, L);
    ldl_solve(n, L, d, b);
}

/* sqp.rs:111-250 solve_qp_active_set and 252-268 feasible_step_size.
 * q is MX x MX row-major. */
static uint64_t qp(int N, const double* q, const double* c, const double* y0, int modify, const ko_tune* t,
                   double* yout) {
    double y[MX];
    int ws[MX];
    for (int i = 0; i < N; i++) {
        y[i] = y0[i];
        ws[i] = (y[i] == 0.0);
    }
    uint64_t iter = 0;
    while (iter < t->qp_max_iter) {
        sum_to_one(N, y, y);
        int fi[MX], fc = 0;
        for (int i = 0; i < N; i++)
            if (!ws[i]) fi[fc++] = i;
        if (fc == 0) abort(); /* sqp.rs:143-144 */
        double yf[MX] = {0}, cf[MX] = {0}, qf[MX * MX] = {0};
        for (int i = 0; i < fc; i++) {
            yf[i] = y[fi[i]];
            cf[i] = c[fi[i]];
            for (int j = 0; j < fc; j++) qf[i * MX + j] = q[fi[i] * MX + fi[j]];
        }
        double gf[MX] = {0};
        for (int i = 0; i < fc; i++) gf[i] = cf[i] + dotn(fc, qf + i * MX, yf);
        double L[MX * MX] = {0}, D[MX] = {0};
        if (modify) gmw(fc, qf, L, D);
        else unmod(fc, qf, L, D);
        ldl_solve(fc, L, D, gf);
        double ones[MX] = {0};
        for (int i = 0; i < fc; i++) ones[i] = 1.0;
        ldl_solve(fc, L, D, ones);
        double sg = 0.0, so = 0.0;
        for (int i = 0; i < fc; i++) sg += gf[i];
        for (int i = 0; i < fc; i++) so += ones[i];
        const double lambda = sg / so;
        for (int i = 0; i < N; i++) gf[i] = lambda * ones[i] - gf[i];
        double m = fabs(gf[0]);
        for (int i = 1; i < N; i++) {
            double v = fabs(gf[i]);
            if (tkey(v) >= tkey(m)) m = v;
        }
        if (m <= t->qp_zero_search_tol) {
            if (fc == N) break;
            double g[MX];
            for (int i = 0; i < N; i++) g[i] = c[i] + dotn(N, q + i * MX, y);
            int idx = -1;
            for (int i = 0; i < N; i++)
                if (ws[i] && (idx < 0 || tkey(g[i]) < tkey(g[idx]))) idx = i;
            const double mult = g[idx] - lambda;
            if (mult >= -t->qp_conv_tol) break;
            ws[idx] = 0;
        } else {
            double p[MX] = {0};
            int k = 0;
            for (int i = 0; i < N; i++)
                if (!ws[i]) p[i] = gf[k++];
            double alpha = 1.0;
            int bi = -1;
            for (int i = 0; i < N; i++)
                if (p[i] < 0.0) {
                    const double a = -y[i] / p[i];
                    if (a < alpha) {
                        alpha = a;
                        bi = i;
                    }
                }
            for (int i = 0; i < N; i++) y[i] = fmax(y[i] + alpha * p[i], 0.0);
            if (bi >= 0) {
                ws[bi] = 1;
                y[bi] = 0.0;
            }
        }
        iter++;
    }
    for (int i = 0; i < N; i++) yout[i] = y[i];
    return iter;
}

uint64_t ko_solve_qp(int N, const double* q, const double* c, const double* y0, int modify, const ko_tune* t,
                     double* y) {
    double Q[MX * MX];
    to_mx(N, q, Q);
    return qp(N, Q, c, y0, modify, t, y);
}

/* sqp.rs:312-391 estimate_condition_number (h is MX x MX) */
static double condition(int N, const double* x, const double* g, const double* h, double tol) {
    const double lambda = dotn(N, x, g);
    int fi[MX], fc = 0;
    for (int i = 0; i < N; i++) {
        int strict = x[i] == 0.0 && g[i] >= lambda + tol;
        if (!strict) fi[fc++] = i;
    }
    if (fc == 0) abort(); /* sqp.rs:336-337 */
    double hf[MX * MX] = {0};
    for (int i = 0; i < fc; i++)
        for (int j = 0; j < fc; j++) hf[i * MX + j] = h[fi[i] * MX + fi[j]];
    double L[MX * MX] = {0}, D[MX] = {0};
    unmod(fc, hf, L, D);
    for (int i = 0; i < fc; i++)
        if (!(D[i] > 0.0)) return INFINITY;
    int imax = 0, imin = 0;
    for (int i = 1; i < fc; i++) {
        if (tkey(D[i]) >= tkey(D[imax])) imax = i; /* max_by: last maximum */
        if (tkey(D[i]) < tkey(D[imin])) imin = i;  /* min_by: first minimum */
    }
    const double dmax = D[imax], dmin = D[imin];
    double w[MX] = {0}, emax[MX] = {0}, emin[MX] = {0};
    emax[imax] = 1.0;
    emin[imin] = 1.0;
    for (int i = 0; i < fc; i++) w[i] = dotn(fc, L + i * MX, emax);
    backsub(fc, L, emin);
    const double nmax = sqrt(dotn(fc, w, w));
    const double nmin = sqrt(dotn(fc, emin, emin));
    return nmax * nmin * (dmax / dmin);
}

double ko_condition(int N, const double* x, const double* g, const double* h, double tol) {
    double H[MX * MX];
    to_mx(N, h, H);
    return condition(N, x, g, H, tol);
}

/* ---------------------------------------------------------------- SQP */

typedef double (*obj_fn)(void* ctx, const double* x, double eps);
typedef void (*gh_fn)(void* ctx, const double* x, double eps, double* g, double* h /* MX x MX */);

/* sqp.rs:95-109 */
static int converged(int N, const double* x, const double* g, double tol) {
    const double lambda = dotn(N, x, g);
    for (int i = 0; i < N; i++) {
        if (x[i] == 0.0) {
            if (!(g[i] >= lambda - tol)) return 0;
        } else if (!(fabs(g[i] - lambda) <= tol)) return 0;
    }
    return 1;
}

/* sqp.rs:49-93 solve_sqp with backtracking_line_search (270-310) */
static void sqp(int N, obj_fn obj, gh_fn gh, void* ctx, const double* x0, const ko_tune* t, double* f_out,
                double* x_out, uint64_t* it_out, double* cond_out) {
    double x[MX];
    sum_to_one(N, x0, x);
    double f = obj(ctx, x, t->epsilon);
    for (uint64_t iter = 0; iter < t->sqp_max_iter; iter++) {
        double g[MX], h[MX * MX];
        gh(ctx, x, t->epsilon, g, h);
        if (converged(N, x, g, t->sqp_conv_tol)) {
            *cond_out = condition(N, x, g, h, t->sqp_conv_tol);
            *f_out = f;
            memcpy(x_out, x, sizeof(double) * N);
            *it_out = iter;
            return;
        }
        double c[MX], y[MX];
        for (int i = 0; i < N; i++) c[i] = g[i] - dotn(N, h + i * MX, x);
        qp(N, h, c, x, 1, t, y);
        double p[MX];
        for (int i = 0; i < N; i++) p[i] = y[i] - x[i];
        const double tt = t->bls_sufficient_decrease * dotn(N, g, p);
        double alpha = 1.0;
        uint64_t bi = 0;
        double xn[MX], fn;
        for (;;) {
            for (int i = 0; i < N; i++) xn[i] = x[i] + alpha * p[i];
            sum_to_one(N, xn, xn);
            fn = obj(ctx, xn, t->epsilon);
            if (fn <= f + alpha * tt) break;
            bi++;
            if (bi == t->bls_max_iter) break;
            alpha *= t->bls_step_size_reduce;
        }
        memcpy(x, xn, sizeof(double) * N);
        f = fn;
    }
    *f_out = f;
    memcpy(x_out, x, sizeof(double) * N);
    *it_out = t->sqp_max_iter;
    *cond_out = NAN;
}

/* ---------------------------------------------------- Jacquard pair fit */

typedef struct {
    int order;
    ko_lanevec lv; /* REF */
    const double* rows; /* SEG */
    const uint8_t* valid;
    const uint32_t* seg;
    size_t nseg;
    const uint8_t *gx, *gy; /* CLASS: the pair's codes and the frequencies */
    const double* af;
    const double* prows;    /* CLASS with tri/tetra-allelic parts: their rows, */
    const uint8_t* pvalid;  /* usable flags and segment bounds, after the */
    const uint32_t* pseg;   /* 2-allele part's class-order sums */
    size_t npseg;
} pair_ctx;

static void class_gh(const pair_ctx* c, const double* x, double eps, double* f, double* g, double* h) {
    double F, G[9], H[45];
    size_t n;
    ko_class_sums(c->gx, c->gy, c->af, c->seg, c->nseg, x, eps, &F, G, H, &n);
    /* the parts' segment partials, folded after the 2-allele part's */
    for (size_t k = 0; k < c->npseg; k++) {
        double zf, zg[9], zh[45];
        size_t zn;
        seg_sums(c->prows, c->pvalid, c->pseg + k, 1, x, eps, 1, 1, &zf, zg, zh, &zn);
        F = F + zf;
        for (int i = 0; i < 9; i++) G[i] = G[i] + zg[i];
        for (int i = 0; i < 45; i++) H[i] = H[i] + zh[i];
        n += zn;
    }
    if (f) *f = -F / (double)n;
    if (g) gh_from_sums(G, H, n, g, h);
}

static double pair_obj(void* vc, const double* x, double eps) {
    pair_ctx* c = (pair_ctx*)vc;
    if (c->order == KO_ORDER_FAST) return ko_jacq_obj_lv_fast(&c->lv, x, eps);
    if (c->order == KO_ORDER_REF) return ko_ops()->jacq_obj(&c->lv, x, eps);
    if (c->order == KO_ORDER_LANES) return ko_jacq_obj_lanes(c->rows, c->valid, 0, c->seg, c->nseg, x, eps);
    if (c->order == KO_ORDER_CLASS) {
        double f;
        class_gh(c, x, eps, &f, NULL, NULL);
        return f;
    }
    return ko_jacq_obj_seg(c->rows, c->valid, 0, c->seg, c->nseg, x, eps);
}

static void pair_gh(void* vc, const double* x, double eps, double* g, double* h) {
    pair_ctx* c = (pair_ctx*)vc;
    double hh[81];
    if (c->order == KO_ORDER_FAST) ko_jacq_gh_lv_fast(&c->lv, x, eps, g, hh);
    else if (c->order == KO_ORDER_REF) ko_ops()->jacq_gh(&c->lv, x, eps, g, hh);
    else if (c->order == KO_ORDER_LANES) ko_jacq_gh_lanes(c->rows, c->valid, 0, c->seg, c->nseg, x, eps, g, hh);
    else if (c->order == KO_ORDER_CLASS) class_gh(c, x, eps, NULL, g, hh);
    else ko_jacq_gh_seg(c->rows, c->valid, 0, c->seg, c->nseg, x, eps, g, hh);
    memcpy(h, hh, sizeof hh);
}

static void fit_ctx(pair_ctx* ctx, int self_pair, const ko_tune* t, ko_output* out) {
    /* coefficients.rs:88-92 */
    double x0[9];
    if (self_pair) {
        for (int i = 0; i < 9; i++) x0[i] = 0.0;
        x0[0] = 0.5;
        x0[6] = 0.5;
    } else
        for (int i = 0; i < 9; i++) x0[i] = 1.0 / 9.0;
    sqp(9, pair_obj, pair_gh, ctx, x0, t, &out->obj, out->jacquard, &out->iters, &out->cond);
}

void ko_fit_pair_rows(int order, const double* rows, size_t n_or_V, const uint8_t* valid, const uint32_t* seg,
                      size_t nseg, int self_pair, const ko_tune* t, ko_output* out) {
    pair_ctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.order = order;
    if (order == KO_ORDER_REF) {
        rows_to_lanevec(rows, n_or_V, &ctx.lv);
        out->n_sites = (uint32_t)n_or_V;
    } else {
        ctx.rows = rows;
        ctx.valid = valid;
        ctx.seg = seg;
        ctx.nseg = nseg;
        uint32_t n = 0;
        for (size_t s = 0; s < n_or_V; s++) n += valid[s] != 0;
        out->n_sites = n;
    }
    fit_ctx(&ctx, self_pair, t, out);
    if (order == KO_ORDER_REF) free(ctx.lv.blocks);
}

static void class_parts_ctx(pair_ctx* ctx, const uint32_t* seg2, const uint8_t* gx, const uint8_t* gy,
                            const double* af2, const double* prows, const uint8_t* pvalid, const uint32_t* pseg,
                            size_t npseg) {
    memset(ctx, 0, sizeof *ctx);
    ctx->order = KO_ORDER_CLASS;
    ctx->gx = gx;
    ctx->gy = gy;
    ctx->af = af2;
    ctx->seg = seg2;
    ctx->nseg = 1;
    ctx->prows = prows;
    ctx->pvalid = pvalid;
    ctx->pseg = pseg;
    ctx->npseg = npseg;
}

void ko_class_parts_eval(const uint8_t* gx, const uint8_t* gy, const double* af2, uint32_t V2, const double* prows,
                         const uint8_t* pvalid, const uint32_t* pseg, size_t npseg, const double x[9], double eps,
                         double* f, double g[9], double h[81]) {
    pair_ctx ctx;
    const uint32_t seg2[2] = {0, V2};
    class_parts_ctx(&ctx, seg2, gx, gy, af2, prows, pvalid, pseg, npseg);
    class_gh(&ctx, x, eps, f, g, h);
}

void ko_fit_pair_class_parts(const uint8_t* gx, const uint8_t* gy, const double* af2, uint32_t V2,
                             const double* prows, const uint8_t* pvalid, const uint32_t* pseg, size_t npseg,
                             int self_pair, const ko_tune* t, ko_output* out) {
    pair_ctx ctx;
    const uint32_t seg2[2] = {0, V2};
    class_parts_ctx(&ctx, seg2, gx, gy, af2, prows, pvalid, pseg, npseg);
    uint32_t n = 0;
    for (uint32_t s = 0; s < V2; s++) n += gx[s] < 3 && gy[s] < 3;
    for (size_t k = 0; k < npseg; k++)
        for (uint32_t s = pseg[k]; s < pseg[k + 1]; s++) n += pvalid[s] != 0;
    out->n_sites = n;
    fit_ctx(&ctx, self_pair, t, out);
}

static const double ONEHOT[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {NAN, NAN, NAN}};

size_t ko_pair_rows_codes(const uint8_t* gx, const uint8_t* gy, const double* af, size_t V, double* rows,
                          uint8_t* valid) {
    size_t n = 0;
    for (size_t s = 0; s < V; s++) {
        double c[9];
        ko_mixture_row2(ONEHOT[gx[s] & 3], ONEHOT[gy[s] & 3], af + 2 * s, c);
        int ok = 1;
        for (int k = 0; k < 9; k++)
            if (c[k] != c[k]) ok = 0; /* lanevector.rs:100-104 */
        valid[s] = (uint8_t)ok;
        for (int k = 0; k < 9; k++) rows[s * 9 + k] = ok ? c[k] : 0.0;
        n += ok;
    }
    return n;
}

/* Soft likelihoods: lx, ly are V x 3 linear-scale likelihood rows (vcf.rs:
 * 10^(gl - max) or 10^(-pl/10)); a row with a NaN drops the site for the pair. */
size_t ko_pair_rows_gl(const double* lx, const double* ly, const double* af, size_t V, double* rows,
                       uint8_t* valid) {
    size_t n = 0;
    for (size_t s = 0; s < V; s++) {
        double c[9];
        ko_mixture_row2(lx + 3 * s, ly + 3 * s, af + 2 * s, c);
        int ok = 1;
        for (int k = 0; k < 9; k++)
            if (c[k] != c[k]) ok = 0; /* lanevector.rs:100-104 */
        valid[s] = (uint8_t)ok;
        for (int k = 0; k < 9; k++) rows[s * 9 + k] = ok ? c[k] : 0.0;
        n += ok;
    }
    return n;
}

/* ------------------------------------------------- allele frequencies */

typedef struct {
    const double* blocks;
    size_t nb;
    const double* rem;
    size_t nrem;
} allele_ctx;

static double allele_obj(void* vc, const double* x, double eps) {
    allele_ctx* c = (allele_ctx*)vc;
    return ko_ops()->allele_obj(c->blocks, c->nb, c->rem, c->nrem, x, eps);
}

static void allele_gh(void* vc, const double* x, double eps, double* g, double* h) {
    allele_ctx* c = (allele_ctx*)vc;
    double hh[4];
    ko_ops()->allele_gh(c->blocks, c->nb, c->rem, c->nrem, x, eps, g, hh);
    h[0] = hh[0];
    h[1] = hh[1];
    h[MX] = hh[2];
    h[MX + 1] = hh[3];
}

/* allele/mod.rs:100-126 for one site: lanevector.rs:43-77 fill (NaN rows
 * skipped), SQP from the uniform start, maf zeroing, renormalise. */
static int allele_site(const double* gl, size_t n, double* blocks, uint64_t iterations, double maf, double af[2],
                       uint64_t* iters) {
    size_t cnt = 0;
    double rem[8 * 3];
    double blk[24];
    size_t nb = 0, cur = 0;
    for (size_t i = 0; i < n; i++) {
        const double* r = gl + 3 * i;
        if (r[0] != r[0] || r[1] != r[1] || r[2] != r[2]) continue;
        for (int k = 0; k < 3; k++) blk[k * 8 + cur] = r[k];
        cur++;
        cnt++;
        if (cur == 8) {
            memcpy(blocks + nb * 24, blk, sizeof blk);
            nb++;
            cur = 0;
        }
    }
    for (size_t r = 0; r < cur; r++)
        for (int k = 0; k < 3; k++) rem[r * 3 + k] = blk[k * 8 + r];
    if (cnt == 0) return -1; /* lanevector.rs:73-76 */
    allele_ctx ctx = {blocks, nb, rem, cur};
    ko_tune t;
    ko_tune_new(&t, iterations, 3); /* allele/mod.rs:93: qp cap A + 1 */
    double x0[2] = {0.5, 0.5}, x[2], f, cond;
    uint64_t it;
    sqp(2, allele_obj, allele_gh, &ctx, x0, &t, &f, x, &it, &cond);
    for (int i = 0; i < 2; i++)
        if (x[i] > 0.0 && x[i] < maf) x[i] = 0.0;
    sum_to_one(2, x, af);
    if (iters) *iters = it;
    return 0;
}

int ko_allele_freq2_rows(const double* gl, size_t n, uint64_t iterations, double maf, double af[2],
                         uint64_t* iters) {
    double* blocks = (double*)malloc(((n / 8) + 1) * 24 * sizeof(double));
    int rc = allele_site(gl, n, blocks, iterations, maf, af, iters);
    free(blocks);
    return rc;
}

/* allele/objective.rs:114-209 and grad_hess.rs:169-396 for A alleles (the
 * lane2/3/4 functions are one pattern): G = A(A+1)/2 likelihoods per row in
 * VCF order, rows as 8-lane blocks [G][8] + remainder [G]. */
static inline double radd8c(const double v[8]) {
    double a[4], b[2];
    for (int i = 0; i < 4; i++) a[i] = v[i] + v[i + 4];
    for (int i = 0; i < 2; i++) b[i] = a[i] + a[i + 2];
    return b[0] + b[1];
}

typedef struct {
    int A, G;
    const double* blocks;
    size_t nb;
    const double* rem;
    size_t nrem;
} allele_ctxA;

static double allele_rowA_obj(const allele_ctxA* c, const double* w, const double* l, size_t stride, double eps) {
    double p = eps;
    for (int g = 0; g < c->G; g++) p = fma(w[g], l[g * stride], p);
    return ko_log(p);
}

static double allele_objA(void* vc, const double* x, double eps) {
    const allele_ctxA* c = (const allele_ctxA*)vc;
    int gi[10], gj[10];
    genotype_alleles(c->A, gi, gj);
    double w[10];
    for (int g = 0; g < c->G; g++) {
        if (gi[g] == gj[g]) w[g] = x[gi[g]] * x[gi[g]];
        else {
            const double t = x[gi[g]] * x[gj[g]];
            w[g] = t + t;
        }
    }
    double zs[8] = {0};
    for (size_t b = 0; b < c->nb; b++)
        for (int j = 0; j < 8; j++) zs[j] = allele_rowA_obj(c, w, c->blocks + b * c->G * 8 + j, 8, eps) + zs[j];
    double rs = 0.0;
    for (size_t r = 0; r < c->nrem; r++) rs = allele_rowA_obj(c, w, c->rem + r * c->G, 1, eps) + rs;
    return -(radd8c(zs) + rs) / (double)(c->nb * 8 + c->nrem);
}

static void allele_rowA_gh(const allele_ctxA* c, const int* gi, const int* gj, const double* x, const double* l,
                           size_t stride, double eps, double* zg, double* zh) {
    const int A = c->A;
    double lx[4], lxd[4];
    for (int a = 0; a < A; a++) {
        /* lx_a = sum_b x_b l_{ab}: first a product, then fma over b */
        for (int b = 0; b < A; b++) {
            const int i = a < b ? a : b, j = a < b ? b : a;
            const double lab = l[(j * (j + 1) / 2 + i) * stride];
            lx[a] = b == 0 ? x[0] * lab : fma(x[b], lab, lx[a]);
        }
    }
    double d = eps;
    for (int a = 0; a < A; a++) d = fma(x[a], lx[a], d);
    const double inv = 1.0 / d;
    for (int a = 0; a < A; a++) lxd[a] = lx[a] * inv;
    for (int g = 0; g < c->G; g++) {
        zh[g] = fma(2.0, lxd[gi[g]] * lxd[gj[g]], zh[g]);
        zh[g] = fma(-l[g * stride], inv, zh[g]);
    }
    for (int a = 0; a < A; a++) zg[a] = fma(lx[a], inv, zg[a]);
}

static void allele_ghA(void* vc, const double* x, double eps, double* g, double* h) {
    const allele_ctxA* c = (const allele_ctxA*)vc;
    int gi[10], gj[10];
    genotype_alleles(c->A, gi, gj);
    double zg[8][4], zh[8][10], rg[4] = {0}, rh[10] = {0};
    memset(zg, 0, sizeof zg);
    memset(zh, 0, sizeof zh);
    for (size_t b = 0; b < c->nb; b++)
        for (int j = 0; j < 8; j++)
            allele_rowA_gh(c, gi, gj, x, c->blocks + b * c->G * 8 + j, 8, eps, zg[j], zh[j]);
    for (size_t r = 0; r < c->nrem; r++) allele_rowA_gh(c, gi, gj, x, c->rem + r * c->G, 1, eps, rg, rh);
    const double n = (double)(c->nb * 8 + c->nrem);
    for (int a = 0; a < c->A; a++) {
        double v[8];
        for (int j = 0; j < 8; j++) v[j] = zg[j][a];
        g[a] = -2.0 * (radd8c(v) + rg[a]) / n;
    }
    for (int gg = 0; gg < c->G; gg++) {
        double v[8];
        for (int j = 0; j < 8; j++) v[j] = zh[j][gg];
        const double val = 2.0 * (radd8c(v) + rh[gg]) / n;
        h[gi[gg] * MX + gj[gg]] = val;
        h[gj[gg] * MX + gi[gg]] = val;
    }
}

int ko_allele_freqA_rows(int A, const double* gl, size_t n, uint64_t iterations, double maf, double* af,
                         uint64_t* iters) {
    const int G = A * (A + 1) / 2;
    double* blocks = (double*)malloc(((n / 8) + 1) * G * 8 * sizeof(double));
    double rem[8 * 10], blk[10 * 8];
    size_t nb = 0, cur = 0, cnt = 0;
    for (size_t i = 0; i < n; i++) {
        const double* r = gl + (size_t)G * i;
        int nan = 0;
        for (int k = 0; k < G; k++) nan |= r[k] != r[k];
        if (nan) continue; /* lanevector.rs:51-53 */
        for (int k = 0; k < G; k++) blk[k * 8 + cur] = r[k];
        cur++;
        cnt++;
        if (cur == 8) {
            memcpy(blocks + nb * G * 8, blk, (size_t)G * 8 * sizeof(double));
            nb++;
            cur = 0;
        }
    }
    for (size_t r = 0; r < cur; r++)
        for (int k = 0; k < G; k++) rem[r * G + k] = blk[k * 8 + r];
    if (cnt == 0) {
        free(blocks);
        return -1; /* lanevector.rs:73-76 */
    }
    allele_ctxA ctx = {A, G, blocks, nb, rem, cur};
    ko_tune t;
    ko_tune_new(&t, iterations, (uint64_t)A + 1); /* allele/mod.rs:93 */
    double x0[4], x[4], f, cond;
    for (int a = 0; a < A; a++) x0[a] = 1.0 / (double)A; /* allele/mod.rs:106 */
    uint64_t it;
    sqp(A, allele_objA, allele_ghA, &ctx, x0, &t, &f, x, &it, &cond);
    for (int a = 0; a < A; a++)
        if (x[a] > 0.0 && x[a] < maf) x[a] = 0.0;
    sum_to_one(A, x, af);
    if (iters) *iters = it;
    free(blocks);
    return 0;
}

/* Pair rows of one pair over sites with A alleles: lx, ly V x G, af V x A. */
size_t ko_pair_rows_generic(int A, const double* lx, const double* ly, const double* af, size_t V, double* rows,
                            uint8_t* valid) {
    const int G = A * (A + 1) / 2;
    size_t n = 0;
    for (size_t s = 0; s < V; s++) {
        double c[9];
        ko_mixture_generic(A, lx + (size_t)G * s, ly + (size_t)G * s, af + (size_t)A * s, c);
        int ok = 1;
        for (int k = 0; k < 9; k++)
            if (c[k] != c[k]) ok = 0; /* lanevector.rs:100-104 */
        valid[s] = (uint8_t)ok;
        for (int k = 0; k < 9; k++) rows[s * 9 + k] = ok ? c[k] : 0.0;
        n += ok;
    }
    return n;
}

typedef struct {
    const uint8_t* codes;
    const double* gl; /* N x V x 3 likelihood rows, instead of codes when set */
    uint32_t N, V;
    uint64_t iterations;
    double maf;
    double* af;
    atomic_size_t next;
    atomic_int err;
} af_job;

static void* af_worker(void* vj) {
    af_job* j = (af_job*)vj;
    double* gl = (double*)malloc((size_t)j->N * 3 * sizeof(double));
    double* blocks = (double*)malloc(((j->N / 8) + 1) * 24 * sizeof(double));
    for (;;) {
        size_t s0 = atomic_fetch_add(&j->next, 256);
        if (s0 >= j->V) break;
        size_t s1 = s0 + 256 < j->V ? s0 + 256 : j->V;
        for (size_t s = s0; s < s1; s++) {
            for (uint32_t i = 0; i < j->N; i++)
                memcpy(gl + 3 * i, j->gl ? j->gl + ((size_t)i * j->V + s) * 3 : ONEHOT[j->codes[(size_t)i * j->V + s] & 3],
                       24);
            if (allele_site(gl, j->N, blocks, j->iterations, j->maf, j->af + 2 * s, NULL)) atomic_store(&j->err, -1);
        }
    }
    free(gl);
    free(blocks);
    return NULL;
}

static int nthreads_default(int n) {
    if (n > 0) return n;
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

static int allele_freqs(const uint8_t* codes, const double* glrows, uint32_t N, uint32_t V, uint64_t iterations,
                        double maf, int nthreads, double* af_out) {
    af_job j;
    j.codes = codes;
    j.gl = glrows;
    j.N = N;
    j.V = V;
    j.iterations = iterations;
    j.maf = maf;
    j.af = af_out;
    atomic_init(&j.next, 0);
    atomic_init(&j.err, 0);
    int nt = nthreads_default(nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, af_worker, &j);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
    return atomic_load(&j.err);
}

int ko_allele_freqs_codes(const uint8_t* codes, uint32_t N, uint32_t V, uint64_t iterations, double maf,
                          int nthreads, double* af_out) {
    return allele_freqs(codes, NULL, N, V, iterations, maf, nthreads, af_out);
}

int ko_allele_freqs_gl(const double* gl, uint32_t N, uint32_t V, uint64_t iterations, double maf, int nthreads,
                       double* af_out) {
    return allele_freqs(NULL, gl, N, V, iterations, maf, nthreads, af_out);
}

/* --------------------------------------------------------- pair loop */

typedef struct {
    int order;
    const uint8_t* codes;
    const double* gl; /* N x V x 3 likelihood rows, instead of codes when set */
    uint32_t N, V;
    const double* af;
    const uint32_t* pairs;
    size_t npairs;
    uint32_t nsel;
    const uint32_t* seg;
    size_t nseg;
    ko_tune tune;
    ko_output* out;
    atomic_size_t next;
    const double *p, *q; /* KO_ORDER_FAST: the frequencies split into two arrays */
} fit_job;

/* row-major upper triangle incl. diagonal: index -> (x, y) */
static void tri_pair(size_t idx, uint32_t n, uint32_t* x, uint32_t* y) {
    uint32_t r = 0;
    size_t base = 0;
    while (base + (n - r) <= idx) {
        base += n - r;
        r++;
    }
    *x = r;
    *y = (uint32_t)(r + (idx - base));
}

static void* fit_worker(void* vj) {
    fit_job* j = (fit_job*)vj;
    const size_t V = j->V;
    double* rows = (double*)malloc(V * 9 * sizeof(double));
    uint8_t* valid = (uint8_t*)malloc(V);
    double* comp = (double*)malloc(V * 9 * sizeof(double));
    pair_ctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.order = j->order;
    ctx.lv.blocks = (double*)malloc(((V / 8) + 1) * 72 * sizeof(double));
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->npairs) break;
        uint32_t x, y;
        if (j->pairs) {
            x = j->pairs[2 * i];
            y = j->pairs[2 * i + 1];
        } else
            tri_pair(i, j->nsel, &x, &y);
        ko_output* o = j->out + i;
        memset(o, 0, sizeof *o);
        o->x = x;
        o->y = y;
        if (j->order == KO_ORDER_FAST) {
            /* the lane vector filled directly (LaneVectorBuilder) */
            const size_t n = ko_pair_lanes_codes_fast(j->codes + (size_t)x * V, j->codes + (size_t)y * V, j->p, j->q,
                                                      V, &ctx.lv);
            o->n_sites = (uint32_t)n;
            if (n == 0) {
                for (int k = 0; k < 9; k++) o->jacquard[k] = NAN;
                o->obj = NAN;
                o->cond = NAN;
                continue;
            }
            fit_ctx(&ctx, x == y, &j->tune, o);
            continue;
        }
        if (j->order == KO_ORDER_CLASS) {
            ctx.gx = j->codes + (size_t)x * V;
            ctx.gy = j->codes + (size_t)y * V;
            ctx.af = j->af;
            ctx.seg = j->seg;
            ctx.nseg = j->nseg;
            uint32_t n = 0;
            for (size_t s = 0; s < V; s++) n += ctx.gx[s] < 3 && ctx.gy[s] < 3;
            o->n_sites = n;
            if (n == 0) {
                for (int k = 0; k < 9; k++) o->jacquard[k] = NAN;
                o->obj = NAN;
                o->cond = NAN;
                continue;
            }
            fit_ctx(&ctx, x == y, &j->tune, o);
            continue;
        }
        size_t n = j->gl ? ko_pair_rows_gl(j->gl + (size_t)x * V * 3, j->gl + (size_t)y * V * 3, j->af, V, rows, valid)
                         : ko_pair_rows_codes(j->codes + (size_t)x * V, j->codes + (size_t)y * V, j->af, V, rows,
                                              valid);
        o->n_sites = (uint32_t)n;
        if (n == 0) { /* lanevector.rs:133-138 would panic; report NaN */
            for (int k = 0; k < 9; k++) o->jacquard[k] = NAN;
            o->obj = NAN;
            o->cond = NAN;
            continue;
        }
        if (j->order == KO_ORDER_REF) {
            /* LaneVectorBuilder (lanevector.rs:99-139): compacted rows in
             * 8-lane blocks + remainder */
            size_t m = 0;
            for (size_t s = 0; s < V; s++)
                if (valid[s]) memcpy(comp + 9 * m++, rows + 9 * s, 72);
            ctx.lv.nblocks = n / 8;
            ctx.lv.nrem = n % 8;
            for (size_t b = 0; b < ctx.lv.nblocks; b++)
                for (int l = 0; l < 8; l++)
                    for (int c = 0; c < 9; c++) ctx.lv.blocks[b * 72 + c * 8 + l] = comp[(b * 8 + l) * 9 + c];
            for (size_t r = 0; r < ctx.lv.nrem; r++)
                for (int c = 0; c < 9; c++) ctx.lv.rem[r * 9 + c] = comp[(ctx.lv.nblocks * 8 + r) * 9 + c];
        } else {
            ctx.rows = rows;
            ctx.valid = valid;
            ctx.seg = j->seg;
            ctx.nseg = j->nseg;
        }
        fit_ctx(&ctx, x == y, &j->tune, o);
    }
    free(rows);
    free(valid);
    free(comp);
    free(ctx.lv.blocks);
    return NULL;
}

static int fit_pairs(int order, const uint8_t* codes, const double* glrows, uint32_t N, uint32_t V,
                     const double* af, const uint32_t* pairs, size_t npairs, uint32_t nsel, const uint32_t* seg,
                     size_t nseg, uint64_t iterations, int nthreads, ko_output* out) {
    if (nsel > N) return -1; /* coefficients.rs:53-55 */
    fit_job j;
    memset(&j, 0, sizeof j);
    j.order = order;
    j.codes = codes;
    j.gl = glrows;
    j.N = N;
    j.V = V;
    j.af = af;
    j.pairs = pairs;
    j.nsel = nsel;
    j.npairs = pairs ? npairs : (size_t)nsel * (nsel + 1) / 2;
    j.seg = seg;
    j.nseg = nseg;
    ko_tune_new(&j.tune, iterations, 10); /* coefficients.rs:73 */
    j.out = out;
    atomic_init(&j.next, 0);
    double* pq = NULL;
    if (order == KO_ORDER_FAST) {
        if (glrows || !ko_has_avx512()) {
            j.order = KO_ORDER_REF;  /* the fast loops cover hard calls on AVX-512 hosts */
        } else {
            pq = (double*)malloc(2 * (size_t)V * sizeof(double) + 64);
            for (uint32_t s = 0; s < V; s++) {
                pq[s] = af[2 * s];
                pq[V + s] = af[2 * s + 1];
            }
            j.p = pq;
            j.q = pq + V;
        }
    }
    int nt = nthreads_default(nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, fit_worker, &j);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
    free(pq);
    return 0;
}

int ko_fit_pairs_codes(int order, const uint8_t* codes, uint32_t N, uint32_t V, const double* af,
                       const uint32_t* pairs, size_t npairs, uint32_t nsel, const uint32_t* seg, size_t nseg,
                       uint64_t iterations, int nthreads, ko_output* out) {
    return fit_pairs(order, codes, NULL, N, V, af, pairs, npairs, nsel, seg, nseg, iterations, nthreads, out);
}

int ko_fit_pairs_gl(int order, const double* gl, uint32_t N, uint32_t V, const double* af, const uint32_t* pairs,
                    size_t npairs, uint32_t nsel, const uint32_t* seg, size_t nseg, uint64_t iterations, int nthreads,
                    ko_output* out) {
    return fit_pairs(order, NULL, gl, N, V, af, pairs, npairs, nsel, seg, nseg, iterations, nthreads, out);
}

/* Pair fits over re-bucketed 2-, 3- and 4-allele sites in the reference's
 * order (coefficients.rs:97-107: the 8-lane parts of the 2-, 3- and 4-allele
 * rows, then their remainders, each compacted to the usable rows by
 * LaneVectorBuilder, lanevector.rs:99-139).  gl_A is sample-major N x V_A x
 * G_A, af_A is V_A x A; the 2-allele rows come from mixture.rs:96-236, the
 * others from mixture.rs:8-60.  Threads take pairs with an atomic counter. */
typedef struct {
    const double* gl[3];
    const double* af[3];
    uint32_t V[3];
    const uint32_t* pairs;
    size_t npairs;
    ko_tune tune;
    ko_output* out;
    atomic_size_t next;
} multi_job;

static void* multi_worker(void* vj) {
    multi_job* j = (multi_job*)vj;
    const size_t Vt = (size_t)j->V[0] + j->V[1] + j->V[2];
    double* rows[3];
    uint8_t* valid[3];
    for (int b = 0; b < 3; b++) {
        rows[b] = (double*)malloc(((size_t)j->V[b] + 1) * 9 * sizeof(double));
        valid[b] = (uint8_t*)malloc((size_t)j->V[b] + 1);
    }
    double* comp = (double*)malloc((Vt + 1) * 9 * sizeof(double));
    pair_ctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.order = KO_ORDER_REF;
    ctx.lv.blocks = (double*)malloc(((Vt / 8) + 1) * 72 * sizeof(double));
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->npairs) break;
        const uint32_t x = j->pairs[2 * i], y = j->pairs[2 * i + 1];
        ko_output* o = j->out + i;
        memset(o, 0, sizeof *o);
        o->x = x;
        o->y = y;
        for (int b = 0; b < 3; b++) {
            const int A = b + 2, G = A * (A + 1) / 2;
            const size_t V = j->V[b];
            if (V == 0) continue;
            const double* lx = j->gl[b] + (size_t)x * V * G;
            const double* ly = j->gl[b] + (size_t)y * V * G;
            if (A == 2)
                ko_pair_rows_gl(lx, ly, j->af[b], V, rows[b], valid[b]);
            else
                ko_pair_rows_generic(A, lx, ly, j->af[b], V, rows[b], valid[b]);
        }
        size_t n = 0;
        for (int half = 0; half < 2; half++)
            for (int b = 0; b < 3; b++) {
                const size_t m = (j->V[b] / 8) * 8;
                const size_t lo = half ? m : 0, hi = half ? j->V[b] : m;
                for (size_t s = lo; s < hi; s++)
                    if (valid[b][s]) memcpy(comp + 9 * n++, rows[b] + 9 * s, 72);
            }
        o->n_sites = (uint32_t)n;
        if (n == 0) { /* lanevector.rs:133-138 would panic; report NaN */
            for (int k = 0; k < 9; k++) o->jacquard[k] = NAN;
            o->obj = NAN;
            o->cond = NAN;
            continue;
        }
        ctx.lv.nblocks = n / 8;
        ctx.lv.nrem = n % 8;
        for (size_t bl = 0; bl < ctx.lv.nblocks; bl++)
            for (int l = 0; l < 8; l++)
                for (int c = 0; c < 9; c++) ctx.lv.blocks[bl * 72 + c * 8 + l] = comp[(bl * 8 + l) * 9 + c];
        for (size_t r = 0; r < ctx.lv.nrem; r++)
            for (int c = 0; c < 9; c++) ctx.lv.rem[r * 9 + c] = comp[(ctx.lv.nblocks * 8 + r) * 9 + c];
        fit_ctx(&ctx, x == y, &j->tune, o);
    }
    for (int b = 0; b < 3; b++) {
        free(rows[b]);
        free(valid[b]);
    }
    free(comp);
    free(ctx.lv.blocks);
    return NULL;
}

int ko_fit_pairs_multi(uint32_t N, const double* gl2, const double* af2, uint32_t V2, const double* gl3,
                       const double* af3, uint32_t V3, const double* gl4, const double* af4, uint32_t V4,
                       const uint32_t* pairs, size_t npairs, uint64_t iterations, int nthreads, ko_output* out) {
    for (size_t i = 0; i < 2 * npairs; i++)
        if (pairs[i] >= N) return -1;
    multi_job j;
    memset(&j, 0, sizeof j);
    j.gl[0] = gl2;
    j.gl[1] = gl3;
    j.gl[2] = gl4;
    j.af[0] = af2;
    j.af[1] = af3;
    j.af[2] = af4;
    j.V[0] = V2;
    j.V[1] = V3;
    j.V[2] = V4;
    j.pairs = pairs;
    j.npairs = npairs;
    ko_tune_new(&j.tune, iterations, 10); /* coefficients.rs:73 */
    j.out = out;
    atomic_init(&j.next, 0);
    int nt = nthreads_default(nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, multi_worker, &j);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
    return 0;
}

/* ------------------------------------------------- CLS (Brier) path, -f GT */

/* wrapper.rs:412-428 calculate_allele_frequencies::<A>: frequencies by
 * counting both alleles of every call (no missing calls on this path:
 * vcf.rs:217-221 drops such sites).  idx N x V genotype indices in VCF order
 * (j(j+1)/2 + i, coefficients.rs:19-38); af V x A. */
void ko_gt_allele_counts_generic(int A, const uint8_t* idx, uint32_t N, uint32_t V, double* af) {
    int gi[10], gj[10];
    genotype_alleles(A, gi, gj);
    for (uint32_t s = 0; s < V; s++) {
        uint64_t f[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < N; i++) {
            const uint8_t g = idx[(size_t)i * V + s];
            f[gi[g]]++;
            f[gj[g]]++;
        }
        const double tot = (double)(2 * (uint64_t)N);
        for (int a = 0; a < A; a++) af[(size_t)s * A + a] = (double)f[a] / tot;
    }
}

void ko_gt_allele_counts(const uint8_t* codes, uint32_t N, uint32_t V, double* af) {
    ko_gt_allele_counts_generic(2, codes, N, V, af);
}

/* cls.rs:6-26 calculate_m_matrices over IIS_A (conditional.rs:10-50): for
 * every (gx, gy) in genotype order the alleles (i, j), (k, l) of the two
 * genotypes, their IBS mode and iis.rs:279-462 conditional_probabilities.
 * af V x A; m[s][gx][gy][9], G = A(A+1)/2. */
void ko_cls_m_generic(int A, const double* af, uint32_t V, double* m) {
    int gi[10], gj[10];
    genotype_alleles(A, gi, gj);
    const int G = A * (A + 1) / 2;
    for (uint32_t s = 0; s < V; s++) {
        const double* p = af + (size_t)s * A;
        for (int gx = 0; gx < G; gx++)
            for (int gy = 0; gy < G; gy++) {
                const int i = gi[gx], j = gj[gx], k = gi[gy], l = gj[gy];
                ko_conditional_probabilities(p[i], p[j], p[k], p[l], ko_iis_mode(i, j, k, l),
                                             m + (((size_t)s * G + gx) * G + gy) * 9);
            }
    }
}

void ko_cls_m_matrices(const double* af, uint32_t V, double* m) { ko_cls_m_generic(2, af, V, m); }

/* coefficients.rs:280-320 compute_quadratic_matrix: every row of every
 * 2-allele site's m (gx, gy order), then of the 3-allele sites, then of the
 * 4-allele sites, pushed into one 8-lane vector (lanevector.rs:87-139; rows
 * holding a NaN are skipped); quadratic.rs:12-33 Gram matrix divided by the
 * number of sites V = V2 + V3 + V4. */
void ko_cls_gram_multi(const double* m2, uint32_t V2, const double* m3, uint32_t V3, const double* m4, uint32_t V4,
                       double* q) {
    const size_t total = (size_t)V2 * 9 + (size_t)V3 * 36 + (size_t)V4 * 100;
    double* rows = (double*)malloc((total ? total : 1) * 9 * sizeof(double));
    size_t n = 0;
    const double* src[3] = {m2, m3, m4};
    const size_t cnt[3] = {(size_t)V2 * 9, (size_t)V3 * 36, (size_t)V4 * 100};
    for (int b = 0; b < 3; b++)
        for (size_t r = 0; r < cnt[b]; r++) {
            const double* row = src[b] + r * 9;
            int nan = 0;
            for (int c = 0; c < 9; c++) nan |= row[c] != row[c];
            if (nan) continue;
            memcpy(rows + n * 9, row, 9 * sizeof(double));
            n++;
        }
    ko_lanevec lv;
    lv.nblocks = n / 8;
    lv.nrem = n % 8;
    lv.blocks = (double*)malloc((lv.nblocks ? lv.nblocks : 1) * 72 * sizeof(double));
    for (size_t b = 0; b < lv.nblocks; b++)
        for (int j = 0; j < 8; j++)
            for (int c = 0; c < 9; c++) lv.blocks[b * 72 + c * 8 + j] = rows[(b * 8 + j) * 9 + c];
    for (size_t r = 0; r < lv.nrem; r++)
        for (int c = 0; c < 9; c++) lv.rem[r * 9 + c] = rows[(lv.nblocks * 8 + r) * 9 + c];
    ko_ops()->gram(&lv, (double)V2 + (double)V3 + (double)V4, q);
    free(lv.blocks);
    free(rows);
}

void ko_cls_gram(const double* m, uint32_t V, double* q) { ko_cls_gram_multi(m, V, NULL, 0, NULL, 0, q); }

/* cls.rs:28-49 calculate_quadratic_c: sum of m_s[gx_s][gy_s] over the sites of
 * one bucket, in site order, from zero. */
static void cls_bucket_sum(const double* m, int G, const uint8_t* gx, const uint8_t* gy, uint32_t V, double c[9]) {
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    for (uint32_t s = 0; s < V; s++) {
        const double* row = m + (((size_t)s * G + gx[s]) * G + gy[s]) * 9;
        for (int k = 0; k < 9; k++) c[k] = c[k] + row[k];
    }
}

/* coefficients.rs:216-276 for one pair: c = (c2 + (c3 + c4)) / -V, QP from 1/9
 * (or the 2-variable (delta1, delta7) QP from (1/2, 1/2) for a self pair);
 * brier, gradient and condition.  q is 9 x 9 row-major; qp cap = iterations.
 * m[b], x[b], y[b], V[b]: the m-matrices and the two genotype-index rows of
 * the 2-, 3- and 4-allele sites. */
static void cls_pair_multi(const double* const m[3], const double* q, const uint8_t* const x[3],
                           const uint8_t* const y[3], const uint32_t V[3], int self_pair, uint64_t iterations,
                           ko_output* out) {
    static const int G[3] = {3, 6, 10};
    double cb[3][9], c[9];
    for (int b = 0; b < 3; b++) cls_bucket_sum(m[b], G[b], x[b], y[b], V[b], cb[b]);
    const uint32_t vt = V[0] + V[1] + V[2];
    const double nv = -(double)vt;
    for (int k = 0; k < 9; k++) c[k] = (cb[0][k] + (cb[1][k] + cb[2][k])) / nv;
    ko_tune t;
    ko_tune_new(&t, iterations, iterations);
    double Q[MX * MX];
    memcpy(Q, q, sizeof Q);
    double delta[9];
    uint64_t iters;
    if (self_pair) {
        double q2[MX * MX] = {0};
        q2[0] = Q[0];
        q2[1] = Q[6];
        q2[MX] = Q[6 * MX];
        q2[MX + 1] = Q[6 * MX + 6];
        double c2[2] = {c[0], c[6]}, y0[2] = {0.5, 0.5}, yv[2];
        iters = qp(2, q2, c2, y0, 1, &t, yv);
        for (int k = 0; k < 9; k++) delta[k] = 0.0;
        delta[0] = yv[0];
        delta[6] = yv[1];
    } else {
        double y0[9];
        for (int k = 0; k < 9; k++) y0[k] = 1.0 / 9.0;
        iters = qp(9, Q, c, y0, 1, &t, delta);
    }
    double qd[9], g[9];
    for (int i = 0; i < 9; i++) qd[i] = dotn(9, Q + i * MX, delta);
    const double brier = dotn(9, delta, qd) + 2.0 * dotn(9, c, delta) + 1.0;
    for (int i = 0; i < 9; i++) g[i] = qd[i] + c[i];
    out->cond = condition(9, delta, g, Q, t.sqp_conv_tol);
    memcpy(out->jacquard, delta, sizeof delta);
    out->obj = brier;
    out->iters = iters;
    out->n_sites = vt;
}

void ko_cls_pair(const double* m, const double* q, const uint8_t* gx, const uint8_t* gy, uint32_t V, int self_pair,
                 uint64_t iterations, ko_output* out) {
    const double* mm[3] = {m, NULL, NULL};
    const uint8_t* xx[3] = {gx, NULL, NULL};
    const uint8_t* yy[3] = {gy, NULL, NULL};
    const uint32_t vv[3] = {V, 0, 0};
    cls_pair_multi(mm, q, xx, yy, vv, self_pair, iterations, out);
}

typedef struct {
    const uint8_t* idx[3];
    uint32_t N, V[3];
    const double* m[3];
    const double* q;
    const uint32_t* pairs;
    size_t npairs;
    uint32_t nsel;
    uint64_t iterations;
    ko_output* out;
    atomic_size_t next;
} cls_job;

static void* cls_worker(void* vj) {
    cls_job* j = (cls_job*)vj;
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->npairs) break;
        uint32_t x, y;
        if (j->pairs) {
            x = j->pairs[2 * i];
            y = j->pairs[2 * i + 1];
        } else
            tri_pair(i, j->nsel, &x, &y);
        ko_output* o = j->out + i;
        memset(o, 0, sizeof *o);
        o->x = x;
        o->y = y;
        const uint8_t* rx[3];
        const uint8_t* ry[3];
        for (int b = 0; b < 3; b++) {
            rx[b] = j->idx[b] ? j->idx[b] + (size_t)x * j->V[b] : NULL;
            ry[b] = j->idx[b] ? j->idx[b] + (size_t)y * j->V[b] : NULL;
        }
        cls_pair_multi(j->m, j->q, rx, ry, j->V, x == y, j->iterations, o);
    }
    return NULL;
}

/* coefficients.rs:166-278 calculate_coefficients_gt with wrapper.rs:22-87:
 * idx2 / idx3 / idx4 are N x V2 / V3 / V4 genotype indices (VCF order, no
 * missing calls) of the 2-, 3- and 4-allele sites; af2 / af3 / af4 are V_A x A
 * frequencies (all NULL = counted per bucket). */
int ko_cls_fit_pairs_multi(uint32_t N, const uint8_t* idx2, uint32_t V2, const uint8_t* idx3, uint32_t V3,
                           const uint8_t* idx4, uint32_t V4, const double* af2, const double* af3, const double* af4,
                           const uint32_t* pairs, size_t npairs, uint32_t nsel, uint64_t iterations, int nthreads,
                           ko_output* out, double* q_out) {
    if (nsel > N) return -1;
    const uint8_t* idx[3] = {idx2, idx3, idx4};
    const uint32_t V[3] = {V2, V3, V4};
    const double* af_in[3] = {af2, af3, af4};
    const int given = af2 || af3 || af4;
    double* af[3];
    double* m[3];
    for (int b = 0; b < 3; b++) {
        const int A = b + 2, G = A * (A + 1) / 2;
        af[b] = (double*)malloc(sizeof(double) * A * (V[b] ? V[b] : 1));
        if (given) {
            if (V[b]) memcpy(af[b], af_in[b], sizeof(double) * A * V[b]);
        } else
            ko_gt_allele_counts_generic(A, idx[b], N, V[b], af[b]);
        m[b] = (double*)malloc(sizeof(double) * 9 * G * G * (V[b] ? V[b] : 1));
        ko_cls_m_generic(A, af[b], V[b], m[b]);
    }
    double q[81];
    ko_cls_gram_multi(m[0], V2, m[1], V3, m[2], V4, q);
    if (q_out) memcpy(q_out, q, sizeof q);
    cls_job j;
    memset(&j, 0, sizeof j);
    for (int b = 0; b < 3; b++) {
        j.idx[b] = idx[b];
        j.V[b] = V[b];
        j.m[b] = m[b];
    }
    j.N = N;
    j.q = q;
    j.pairs = pairs;
    j.nsel = nsel;
    j.npairs = pairs ? npairs : (size_t)nsel * (nsel + 1) / 2;
    j.iterations = iterations;
    j.out = out;
    atomic_init(&j.next, 0);
    int nt = nthreads_default(nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, cls_worker, &j);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
    for (int b = 0; b < 3; b++) {
        free(m[b]);
        free(af[b]);
    }
    return 0;
}

/* The biallelic form: codes N x V (0, 1, 2 = genotype index), af V x 2 or
 * NULL (counted). */
int ko_cls_fit_pairs(const uint8_t* codes, uint32_t N, uint32_t V, const double* af_in, const uint32_t* pairs,
                     size_t npairs, uint32_t nsel, uint64_t iterations, int nthreads, ko_output* out, double* q_out) {
    return ko_cls_fit_pairs_multi(N, codes, V, NULL, 0, NULL, 0, af_in, NULL, NULL, pairs, npairs, nsel, iterations,
                                  nthreads, out, q_out);
}
