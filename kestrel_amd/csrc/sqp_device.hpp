// sqp_device.hpp -- FP64 device routines of the per-pair solver: the SLEEF-u10
// natural log, simplex helpers, GMW81 / plain LDL^T on n <= 9, the primal
// active-set QP, the convergence test and the condition estimate.  Every
// operation keeps the reference's evaluation order (explicit fma() exactly
// where the reference writes mul_add/.fma(); this file is compiled with
// -ffp-contract=off), so a device result can be compared bit for bit with a
// CPU evaluation in the same summation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kdev {

constexpr int MX = 9;

struct Tune {
    uint64_t sqp_max_iter;
    double sqp_conv_tol;
    uint64_t qp_max_iter;
    double qp_conv_tol;
    double qp_zero_search_tol;
    uint64_t bls_max_iter;
    double bls_sufficient_decrease;
    double bls_step_size_reduce;
    double epsilon;
    // > 0 (mixed-precision form): a converged pair with a held variable whose
    // strict-complementarity test g_i >= lambda + tol is decided within this
    // margin reports -condition, so that the host estimates it again from an
    // FP64 evaluation (the FP32 reciprocal moves g by up to ~1e-7)
    double recond_margin;
};

// log/scalar.rs:17-70 (identical to log/avx512.rs:15-79 on positive normals):
// d = m 2^e, m in [0.75, 1.5); x = (m-1)/(m+1) as a double-word (div_sd,
// log/mod.rs:72-85); log d = e ln2 + 2x + x^3 P(x^2) (poly7, log/mod.rs:87-100)
// summed in double-word arithmetic (log/mod.rs:31-70).
__device__ __forceinline__ double sleef_log(double d0) {
    const bool o = d0 < 2.2250738585072014e-308;
    const double d = o ? d0 * (4294967296.0 * 4294967296.0) : d0;
    const long long bs = __double_as_longlong(d * (1.0 / 0.75));
    long long e = ((bs >> 52) & 0x7ff) - 0x3ff;
    const double m = __longlong_as_double(__double_as_longlong(d) - (e << 52));
    if (o) e -= 64;
    const double u = m - 1.0;
    const double l0 = 1.0 + m;
    const double lz = l0 - 1.0;
    const double l1 = m - lz;
    const double r = 1.0 / l0;
    const double q0 = u * r;
    const double uu = fma(u, r, -q0);
    const double ee = fma(-l0, r, 1.0);
    double q1 = fma(-l1, r, ee);
    q1 = fma(q0, q1, uu);
    const double x2 = q0 * q0;
    const double x4 = x2 * x2;
    const double x8 = x4 * x4;
    const double p3 = fma(x4, 0.1532076988502701353, fma(x2, 0.1525629051003428716, 0.1818605932937785996));
    const double p4 = fma(x4, fma(x2, 0.2222214519839380009, 0.2857142932794299317),
                          fma(x2, 0.399999999963525199, 0.6666666666667333541));
    const double t = fma(x8, p3, p4);
    const double ef = (double)e;
    const double HI = 0.693147180559945286226764, LO = 2.319046813846299558417771e-17;
    const double sp = HI * ef;
    const double sr = fma(HI, ef, -sp);
    const double se = fma(LO, ef, sr);
    const double bx0 = q0 + q0, bx1 = q1 + q1;
    const double s1 = sp + bx0;
    const double z1 = s1 - sp;
    const double t1 = bx0 - z1;
    const double e1 = t1 + (se + bx1);
    const double bb = (x2 * q0) * t;
    const double s2 = s1 + bb;
    const double z2 = s2 - s1;
    const double t2 = bb - z2;
    const double e2 = t2 + e1;
    double res = s2 + e2;
    if (d0 != d0 || d0 < 0.0) res = __longlong_as_double(0x7ff8000000000000LL);
    if (d0 == __longlong_as_double(0x7ff0000000000000LL)) res = d0;
    if (d0 == 0.0) res = -__longlong_as_double(0x7ff0000000000000LL);
    return res;
}

// Correctly rounded 1/x for positive normal x whose reciprocal is normal:
// the IEEE division sequence (two Newton steps on v_rcp_f64, then the
// residual correction q + (1 - x q) q) without the range scaling and
// special-case fixup, which are identities on this domain.
__device__ __forceinline__ double recip_rn(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(e, r, r);
}

// sleef_log restricted to positive normal finite d (the sweep's L = eps + c.D
// >= 1e-8): the same operations without the subnormal scaling and the
// special-case selects, which cannot trigger there.
__device__ __forceinline__ double sleef_log_pos(double d) {
    const int hi = __double2hiint(d * (1.0 / 0.75));
    const int e = ((hi >> 20) & 0x7ff) - 0x3ff;
    const double m = __hiloint2double(__double2hiint(d) - (e << 20), __double2loint(d));
    const double u = m - 1.0;
    const double l0 = 1.0 + m;
    const double lz = l0 - 1.0;
    const double l1 = m - lz;
    const double r = recip_rn(l0);
    const double q0 = u * r;
    const double uu = fma(u, r, -q0);
    const double ee = fma(-l0, r, 1.0);
    double q1 = fma(-l1, r, ee);
    q1 = fma(q0, q1, uu);
    const double x2 = q0 * q0;
    const double x4 = x2 * x2;
    const double x8 = x4 * x4;
    const double p3 = fma(x4, 0.1532076988502701353, fma(x2, 0.1525629051003428716, 0.1818605932937785996));
    const double p4 = fma(x4, fma(x2, 0.2222214519839380009, 0.2857142932794299317),
                          fma(x2, 0.399999999963525199, 0.6666666666667333541));
    const double t = fma(x8, p3, p4);
    const double ef = (double)e;
    const double HI = 0.693147180559945286226764, LO = 2.319046813846299558417771e-17;
    const double sp = HI * ef;
    const double sr = fma(HI, ef, -sp);
    const double se = fma(LO, ef, sr);
    const double bx0 = q0 + q0, bx1 = q1 + q1;
    const double s1 = sp + bx0;
    const double z1 = s1 - sp;
    const double t1 = bx0 - z1;
    const double e1 = t1 + (se + bx1);
    const double bb = (x2 * q0) * t;
    const double s2 = s1 + bb;
    const double z2 = s2 - s1;
    const double t2 = bb - z2;
    const double e2 = t2 + e1;
    return s2 + e2;
}

// algebra.rs:4-11
__device__ __forceinline__ double dotn(int n, const double* x, const double* y) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fma(x[i], y[i], s);
    return s;
}

// algebra.rs:81-84
template <int N>
__device__ __forceinline__ void sum_to_one(const double* x, double* out) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) s += x[i];
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = x[i] / s;
}

// Rust f64::total_cmp key
__device__ __forceinline__ long long tkey(double v) {
    long long k = __double_as_longlong(v);
    return k ^ (long long)(((unsigned long long)(k >> 63)) >> 1);
}

// cholesky.rs:6-65 (a, l row stride MX)
__device__ inline void gmw(int n, const double* a, double* l, double* d) {
    double gamma = 0.0;
    for (int i = 0; i < n; i++) gamma = fmax(gamma, fabs(a[i * MX + i]));
    double xi = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) xi = fmax(xi, fabs(a[i * MX + j]));
    const double eps = 2.220446049250313e-16;
    const double delta = eps * fmax(1.0, gamma + xi);
    const double t = sqrt((double)n * (double)n - 1.0);
    const double ratio = n <= 1 ? 0.0 : xi / t;
    const double beta2 = fmax(fmax(gamma, ratio), eps);
    double p[MX];
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

// cholesky.rs:72-102
__device__ inline void unmodified(int n, const double* a, double* l, double* d) {
    double p[MX];
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
            l[i * MX + j] = (a[i * MX + j] - sij) / dj;
        }
        d[j] = dj;
    }
}

// cholesky.rs:131-140
__device__ inline void backward_sub(int n, const double* l, double* b) {
    for (int i = n - 1; i >= 0; i--) {
        double s = 0.0;
        for (int j = i + 1; j < n; j++) s = fma(b[j], l[j * MX + i], s);
        b[i] -= s;
    }
}

// cholesky.rs:105-128
__device__ inline void ldl_solve(int n, const double* l, const double* d, double* b) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < i; j++) s = fma(b[j], l[i * MX + j], s);
        b[i] -= s;
    }
    for (int i = 0; i < n; i++) b[i] /= d[i];
    backward_sub(n, l, b);
}

// sqp.rs:111-250 (+ feasible_step_size 252-268): minimise 1/2 y'Qy + c'y over
// the simplex from y0 with a primal active set; q has row stride MX.
template <int N>
__device__ inline uint64_t solve_qp(const double* q, const double* c, const double* y0, const Tune& t, double* yout) {
    double y[MX];
    bool ws[MX];
    for (int i = 0; i < N; i++) {
        y[i] = y0[i];
        ws[i] = (y[i] == 0.0);
    }
    uint64_t iter = 0;
    while (iter < t.qp_max_iter) {
        sum_to_one<N>(y, y);
        int fi[MX], fc = 0;
        for (int i = 0; i < N; i++)
            if (!ws[i]) fi[fc++] = i;
        if (fc == 0) break;  // sqp.rs:143-144 asserts this never happens
        double yf[MX], cf[MX], qf[MX * MX];
        for (int i = 0; i < fc; i++) {
            yf[i] = y[fi[i]];
            cf[i] = c[fi[i]];
            for (int j = 0; j < fc; j++) qf[i * MX + j] = q[fi[i] * MX + fi[j]];
        }
        double gf[MX], ones[MX];
        for (int i = 0; i < N; i++) {
            gf[i] = 0.0;
            ones[i] = 0.0;
        }
        for (int i = 0; i < fc; i++) gf[i] = cf[i] + dotn(fc, qf + i * MX, yf);
        double L[MX * MX], D[MX];
        for (int i = 0; i < MX * MX; i++) L[i] = 0.0;
        for (int i = 0; i < MX; i++) D[i] = 0.0;
        gmw(fc, qf, L, D);
        ldl_solve(fc, L, D, gf);
        for (int i = 0; i < fc; i++) ones[i] = 1.0;
        ldl_solve(fc, L, D, ones);
        double sg = 0.0, so = 0.0;
        for (int i = 0; i < fc; i++) sg += gf[i];
        for (int i = 0; i < fc; i++) so += ones[i];
        const double lambda = sg / so;
        for (int i = 0; i < N; i++) gf[i] = lambda * ones[i] - gf[i];
        double m = fabs(gf[0]);
        for (int i = 1; i < N; i++) {
            const double v = fabs(gf[i]);
            if (tkey(v) >= tkey(m)) m = v;
        }
        if (m <= t.qp_zero_search_tol) {
            if (fc == N) break;
            double g[MX];
            for (int i = 0; i < N; i++) g[i] = c[i] + dotn(N, q + i * MX, y);
            int idx = -1;
            for (int i = 0; i < N; i++)
                if (ws[i] && (idx < 0 || tkey(g[i]) < tkey(g[idx]))) idx = i;
            const double mult = g[idx] - lambda;
            if (mult >= -t.qp_conv_tol) break;
            ws[idx] = false;
        } else {
            double p[MX];
            int k = 0;
            for (int i = 0; i < N; i++) p[i] = ws[i] ? 0.0 : gf[k++];
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
                ws[bi] = true;
                y[bi] = 0.0;
            }
        }
        iter++;
    }
    for (int i = 0; i < N; i++) yout[i] = y[i];
    return iter;
}

// sqp.rs:95-109
template <int N>
__device__ inline bool converged(const double* x, const double* g, double tol) {
    const double lambda = dotn(N, x, g);
    bool ok = true;
    for (int i = 0; i < N; i++) {
        if (x[i] == 0.0) ok = ok && (g[i] >= lambda - tol);
        else ok = ok && (fabs(g[i] - lambda) <= tol);
    }
    return ok;
}

// sqp.rs:312-391 (Practical Optimization 8.3.3.2); h has row stride MX.
template <int N>
__device__ inline double condition(const double* x, const double* g, const double* h, double tol) {
    const double lambda = dotn(N, x, g);
    int fi[MX], fc = 0;
    for (int i = 0; i < N; i++) {
        const bool strict = x[i] == 0.0 && g[i] >= lambda + tol;
        if (!strict) fi[fc++] = i;
    }
    const double inf = __longlong_as_double(0x7ff0000000000000LL);
    if (fc == 0) return inf;  // sqp.rs:336-337 asserts this never happens
    double hf[MX * MX], L[MX * MX], D[MX];
    for (int i = 0; i < MX * MX; i++) L[i] = 0.0;
    for (int i = 0; i < fc; i++)
        for (int j = 0; j < fc; j++) hf[i * MX + j] = h[fi[i] * MX + fi[j]];
    unmodified(fc, hf, L, D);
    for (int i = 0; i < fc; i++)
        if (!(D[i] > 0.0)) return inf;
    int imax = 0, imin = 0;
    for (int i = 1; i < fc; i++) {
        if (tkey(D[i]) >= tkey(D[imax])) imax = i;  // max_by keeps the last maximum
        if (tkey(D[i]) < tkey(D[imin])) imin = i;   // min_by keeps the first minimum
    }
    double w[MX], emax[MX], emin[MX];
    for (int i = 0; i < MX; i++) {
        emax[i] = 0.0;
        emin[i] = 0.0;
    }
    emax[imax] = 1.0;
    emin[imin] = 1.0;
    for (int i = 0; i < fc; i++) w[i] = dotn(fc, L + i * MX, emax);
    backward_sub(fc, L, emin);
    const double nmax = sqrt(dotn(fc, w, w));
    const double nmin = sqrt(dotn(fc, emin, emin));
    return nmax * nmin * (D[imax] / D[imin]);
}

// ---------------------------------------------------------------------------
// Register-resident forms of the QP and the condition estimate.
//
// The reference gathers the free variables into a compacted n x n system
// (sqp.rs:134-159, 324-349).  Here the free set is a bit mask and the
// matrices keep their full index space with the rows and columns of held
// variables set to zero; every loop runs over all indices in increasing order
// with static bounds, so the arrays live in registers.  A masked index only
// ever adds fma(x, 0, s) == s (s is never -0 here) or leaves a zero in place,
// so each result equals the compacted computation bit for bit.

template <int N>
__device__ __forceinline__ void gmw_masked(uint32_t fm, const double (&A)[N][N], double (&L)[N][N], double (&D)[N]) {
    double gamma = 0.0, xi = 0.0;
    int n = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        n += (fm >> i) & 1;
        gamma = fmax(gamma, fabs(A[i][i]));
#pragma unroll
        for (int j = i + 1; j < N; j++) xi = fmax(xi, fabs(A[i][j]));
    }
    const double eps = 2.220446049250313e-16;
    const double delta = eps * fmax(1.0, gamma + xi);
    const double t = sqrt((double)n * (double)n - 1.0);
    const double ratio = n <= 1 ? 0.0 : xi / t;
    const double beta2 = fmax(fmax(gamma, ratio), eps);
#pragma unroll
    for (int i = 0; i < N; i++) {
        D[i] = 1.0;
#pragma unroll
        for (int j = 0; j < N; j++) L[i][j] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        if (!((fm >> j) & 1)) continue;
        double pj[N];
        double sjj = 0.0;
#pragma unroll
        for (int s = 0; s < j; s++) {
            pj[s] = D[s] * L[j][s];
            sjj = fma(pj[s], L[j][s], sjj);
        }
        const double cjj = A[j][j] - sjj;
        L[j][j] = 1.0;
        double theta = 0.0;
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            double sij = 0.0;
#pragma unroll
            for (int s = 0; s < j; s++) sij = fma(pj[s], L[i][s], sij);
            const double cij = A[i][j] - sij;
            theta = fmax(theta, fabs(cij));
            L[i][j] = cij;
        }
        const double dj = fmax(fmax(fabs(cjj), theta * theta / beta2), delta);
#pragma unroll
        for (int i = j + 1; i < N; i++) L[i][j] = L[i][j] / dj;
        D[j] = dj;
    }
}

template <int N>
__device__ __forceinline__ void unmodified_masked(uint32_t fm, const double (&A)[N][N], double (&L)[N][N],
                                                  double (&D)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        D[i] = 1.0;
#pragma unroll
        for (int j = 0; j < N; j++) L[i][j] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        if (!((fm >> j) & 1)) continue;
        L[j][j] = 1.0;
        double pj[N];
        double sjj = 0.0;
#pragma unroll
        for (int s = 0; s < j; s++) {
            pj[s] = D[s] * L[j][s];
            sjj = fma(pj[s], L[j][s], sjj);
        }
        const double dj = A[j][j] - sjj;
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            double sij = 0.0;
#pragma unroll
            for (int s = 0; s < j; s++) sij = fma(pj[s], L[i][s], sij);
            const double v = (A[i][j] - sij) / dj;
            L[i][j] = ((fm >> i) & 1) ? v : 0.0;
        }
        D[j] = dj;
    }
}

template <int N>
__device__ __forceinline__ void backward_masked(const double (&L)[N][N], double (&b)[N]) {
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        double s = 0.0;
#pragma unroll
        for (int j = i + 1; j < N; j++) s = fma(b[j], L[j][i], s);
        b[i] -= s;
    }
}

template <int N>
__device__ __forceinline__ void ldl_solve_masked(const double (&L)[N][N], const double (&D)[N], double (&b)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < i; j++) s = fma(b[j], L[i][j], s);
        b[i] -= s;
    }
#pragma unroll
    for (int i = 0; i < N; i++) b[i] /= D[i];
    backward_masked<N>(L, b);
}

// sqp.rs:111-268, register-resident (see above).  q is N x N.
template <int N>
__device__ inline uint64_t solve_qp_reg(const double (&Q)[N][N], const double (&c)[N], const double (&y0)[N],
                                        const Tune& t, double (&y)[N]) {
    uint32_t ws = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        y[i] = y0[i];
        if (y[i] == 0.0) ws |= 1u << i;
    }
    const uint32_t all = (1u << N) - 1u;
    uint64_t iter = 0;
    while (iter < t.qp_max_iter) {
        {
            double sy = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) sy += y[i];
#pragma unroll
            for (int i = 0; i < N; i++) y[i] = y[i] / sy;
        }
        const uint32_t fm = all & ~ws;
        if (fm == 0) break;  // sqp.rs:143-144 asserts this never happens
        double A[N][N], yf[N], gf[N], ones[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const bool fi = (fm >> i) & 1;
            yf[i] = fi ? y[i] : 0.0;
            ones[i] = fi ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < N; j++) A[i][j] = (fi && ((fm >> j) & 1)) ? Q[i][j] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N; j++) s = fma(A[i][j], yf[j], s);
            gf[i] = ((fm >> i) & 1) ? c[i] + s : 0.0;
        }
        double L[N][N], D[N];
        gmw_masked<N>(fm, A, L, D);
        ldl_solve_masked<N>(L, D, gf);
        ldl_solve_masked<N>(L, D, ones);
        double sg = 0.0, so = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) sg += gf[i];
#pragma unroll
        for (int i = 0; i < N; i++) so += ones[i];
        const double lambda = sg / so;
        double m = 0.0;
        long long mk = tkey(0.0);
#pragma unroll
        for (int i = 0; i < N; i++) {
            gf[i] = ((fm >> i) & 1) ? lambda * ones[i] - gf[i] : 0.0;
            const double v = fabs(gf[i]);
            if (tkey(v) >= mk) {
                m = v;
                mk = tkey(v);
            }
        }
        if (m <= t.qp_zero_search_tol) {
            if (fm == all) break;
            double g[N];
#pragma unroll
            for (int i = 0; i < N; i++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < N; j++) s = fma(Q[i][j], y[j], s);
                g[i] = c[i] + s;
            }
            int idx = -1;
            double gbest = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++)
                if (((ws >> i) & 1) && (idx < 0 || tkey(g[i]) < tkey(gbest))) {
                    idx = i;
                    gbest = g[i];
                }
            if (gbest - lambda >= -t.qp_conv_tol) break;
            ws &= ~(1u << idx);
        } else {
            double alpha = 1.0;
            int bi = -1;
#pragma unroll
            for (int i = 0; i < N; i++)
                if (gf[i] < 0.0) {
                    const double a = -y[i] / gf[i];
                    if (a < alpha) {
                        alpha = a;
                        bi = i;
                    }
                }
#pragma unroll
            for (int i = 0; i < N; i++) y[i] = fmax(y[i] + alpha * gf[i], 0.0);
#pragma unroll
            for (int i = 0; i < N; i++)
                if (i == bi) {
                    ws |= 1u << i;
                    y[i] = 0.0;
                }
        }
        iter++;
    }
    return iter;
}

// Whether the strict-complementarity decision of the condition estimate
// (sqp.rs:324-331) is within `margin` of flipping for some held variable.
template <int N>
__device__ inline bool complementarity_marginal(const double (&x)[N], const double (&g)[N], double tol,
                                                double margin) {
    double lambda = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) lambda = fma(x[i], g[i], lambda);
    bool m = false;
#pragma unroll
    for (int i = 0; i < N; i++) m = m || (x[i] == 0.0 && fabs(g[i] - (lambda + tol)) < margin);
    return m;
}

// sqp.rs:312-391, register-resident (see above).
template <int N>
__device__ inline double condition_reg(const double (&x)[N], const double (&g)[N], const double (&H)[N][N], double tol) {
    double lambda = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) lambda = fma(x[i], g[i], lambda);
    uint32_t fm = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        if (!(x[i] == 0.0 && g[i] >= lambda + tol)) fm |= 1u << i;
    const double inf = __longlong_as_double(0x7ff0000000000000LL);
    if (fm == 0) return inf;  // sqp.rs:336-337 asserts this never happens
    double A[N][N], L[N][N], D[N];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) A[i][j] = (((fm >> i) & 1) && ((fm >> j) & 1)) ? H[i][j] : 0.0;
    unmodified_masked<N>(fm, A, L, D);
    bool pos = true;
    int imax = -1, imin = -1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (!((fm >> i) & 1)) continue;
        pos = pos && (D[i] > 0.0);
        if (imax < 0 || tkey(D[i]) >= tkey(D[imax])) imax = i;  // max_by: last maximum
        if (imin < 0 || tkey(D[i]) < tkey(D[imin])) imin = i;   // min_by: first minimum
    }
    if (!pos) return inf;
    double w[N], emin[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) s = fma(L[i][j], j == imax ? 1.0 : 0.0, s);
        w[i] = s;
        emin[i] = i == imin ? 1.0 : 0.0;
    }
    backward_masked<N>(L, emin);
    double nw = 0.0, ne = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) nw = fma(w[i], w[i], nw);
#pragma unroll
    for (int i = 0; i < N; i++) ne = fma(emin[i], emin[i], ne);
    double dmax = 0.0, dmin = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (i == imax) dmax = D[i];
        if (i == imin) dmin = D[i];
    }
    return sqrt(nw) * sqrt(ne) * (dmax / dmin);
}

}  // namespace kdev
