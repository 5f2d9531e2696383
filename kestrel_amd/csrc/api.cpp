// api.cpp -- host side of libkestrel_amd: cohort upload, allele frequencies,
// kept-site repacking, and the sweep/step loop of the pair fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

namespace {

thread_local std::string g_err;

}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const std::string& last_error_message() { return g_err; }

kdev::Tune to_dev(const kestrel_tune& t) {
    kdev::Tune d;
    d.sqp_max_iter = t.sqp_max_iter;
    d.sqp_conv_tol = t.sqp_conv_tol;
    d.qp_max_iter = t.qp_max_iter;
    d.qp_conv_tol = t.qp_conv_tol;
    d.qp_zero_search_tol = t.qp_zero_search_tol;
    d.bls_max_iter = t.bls_max_iter;
    d.bls_sufficient_decrease = t.bls_sufficient_decrease;
    d.bls_step_size_reduce = t.bls_step_size_reduce;
    d.epsilon = t.epsilon;
    d.recond_margin = 0.0;
    return d;
}

namespace {

// Repack the kept sites of the uploaded 2-bit rows into c->geno (rows of
// v_pad / 16 words, padding code 3) and set n_kept / v_pad / row_words.
int repack_kept(kestrel_cohort* c, const uint8_t* dpacked, size_t row_stride_bytes, const std::vector<uint32_t>& kept) {
    c->n_kept = (uint32_t)kept.size();
    c->v_pad = std::max<uint32_t>(SITE_PAD, (c->n_kept + SITE_PAD - 1) / SITE_PAD * SITE_PAD);
    c->row_words = c->v_pad / 16;
    DevBuf<uint32_t> dkept;
    HIPCHK(dkept.alloc(std::max<size_t>(kept.size(), 1)));
    if (!kept.empty())
        HIPCHK(hipMemcpyAsync(dkept.p, kept.data(), kept.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              c->stream));
    HIPCHK(c->geno.alloc((size_t)c->n_samples * c->row_words));
    HIPCHK(klaunch::gather(dpacked, row_stride_bytes, dkept.p, c->n_kept, c->n_samples, c->geno.p, c->row_words,
                           c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return KESTREL_OK;
}

// One evaluation pass over every part of the cohort: the main part's segments,
// then those of the tri/tetra-allelic parts, into consecutive partials.
int launch_sweeps(kestrel_cohort* c, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* active,
                  uint32_t ns, double eps, double* partial, hipStream_t st, bool skip_masked = false,
                  bool lanes = false, ClassLists* cls = nullptr, bool build = true) {
    size_t base;  // the tri/tetra-allelic parts' first partial segment
    if (cls) {    // the class-sorted sweep (c->class_form(); partials in c->nseg_cls segments)
        // (without build the lists hold every slot: cls->cap >= ns)
        for (uint32_t s0 = 0; s0 < ns; s0 += cls->cap) {
            const uint32_t nc = std::min(cls->cap, ns - s0);
            if (build)
                HIPCHK(klaunch::class_build(c->geno.p, c->row_words, sx, sy, active, ns, s0, nc, cls->nblk,
                                            cls->lists.p, cls->counts.p, cls->built.p, st));
            HIPCHK(klaunch::sweep_class(c->af.p, c->v_pad, c->seg_cls.p, c->nseg_cls, cls->lists.p, cls->counts.p,
                                        cls->nblk, pt, active, ns, s0, nc, eps, partial, st,
                                        (int)c->opt_class_waves));
        }
        base = c->nseg_cls;
    } else if (lanes) {  // FP64 fits of 2-bit and biallelic likelihood-row cohorts (checked by the caller)
        HIPCHK(klaunch::sweep_lanes(c->geno.p, c->row_words, c->af.p, c->seg.p, c->nseg, sx, sy, pt, active, ns, eps,
                                    partial, st, c->soft ? c->glrows.p : nullptr, c->v_pad, c->gl_stride));
        return KESTREL_OK;
    } else {
        HIPCHK(klaunch::sweep(c->geno.p, c->row_words, c->af.p, c->seg.p, c->nseg, sx, sy, pt, active, ns, eps,
                              partial, st, c->soft ? c->glrows.p : nullptr, c->v_pad, c->gl_stride, skip_masked,
                              c->mixed));
        base = c->nseg;
    }
    for (const auto& p : c->parts) {
        HIPCHK(klaunch::sweep_multi(p->alleles, p->mask.p, p->row_words, p->rows.p, p->v_pad, p->table.p, p->af.p,
                                    p->seg.p, p->nseg, sx, sy, pt, active, ns, eps, partial + base * NPART * ns, st,
                                    c->mixed ? p->table32.p : nullptr, p->codes.n ? p->codes.p : nullptr, p->lut.p));
        base += p->nseg;
    }
    return KESTREL_OK;
}

// A likelihood-row cohort whose every called row is one-hot (an exact 1.0
// beside two zeros: hard calls fed as likelihoods) is held as its 2-bit codes
// alone -- the mask rows already are those codes, 3 for a missing call -- and
// fitted by the 2-bit sweeps: c_s of a one-hot row is m(gx, gy) bit for bit
// (mixture.rs:96-236 adds exact zeros), so every record is the same.
int codes_if_one_hot(kestrel_cohort* c) {
    if (!c->soft || c->gl_stride != 2 || c->n_kept == 0) return KESTREL_OK;
    DevBuf<int> nz;
    HIPCHK(nz.alloc(1));
    HIPCHK(hipMemsetAsync(nz.p, 0, sizeof(int), c->stream));
    HIPCHK(klaunch::any_nonzero(c->glrows.p, c->glrows.n, nz.p, c->stream));
    int any = 0;
    HIPCHK(hipMemcpyAsync(&any, nz.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (any) return KESTREL_OK;
    c->soft = false;
    c->glrows.release();
    DevBuf<unsigned long long> dmiss;
    HIPCHK(dmiss.alloc(1));
    HIPCHK(hipMemsetAsync(dmiss.p, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(klaunch::count_missing(c->geno.p, c->n_samples, c->row_words, c->n_kept, dmiss.p, c->stream));
    unsigned long long miss = 0;
    HIPCHK(hipMemcpyAsync(&miss, dmiss.p, sizeof(miss), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->missing_frac = (double)miss / ((double)c->n_samples * (double)c->n_kept);
    return KESTREL_OK;
}

// Largest likelihood of n values (NaN marks a missing call); a negative or
// infinite likelihood is rejected (vcf.rs:368, 407 make them 10^x >= 0).
int scan_likelihoods(const double* v, size_t n, double* lmax) {
    double m = 0.0;
    bool bad = false;
    for (size_t i = 0; i < n; i++) {
        const double x = v[i];
        bad |= x < 0.0 || x == INFINITY;
        m = x > m ? x : m;  // NaN compares false
    }
    if (bad) return fail(KESTREL_EINVAL, "likelihoods must be non-negative and finite (NaN marks a missing call)");
    *lmax = std::max(*lmax, m);
    return KESTREL_OK;
}

}  // namespace

int launch_exact(kestrel_cohort* c, const uint32_t* sx, const uint32_t* sy, const double* pt, const int* exact,
                 uint32_t ns, double eps, double* partial, hipStream_t st) {
    const int kind = c->soft ? (c->gl_stride == 2 ? 2 : 3) : 0;
    HIPCHK(klaunch::sweep_exact(kind, c->geno.p, c->row_words, c->soft ? c->glrows.p : nullptr, c->v_pad, c->af.p,
                                c->seg.p, c->nseg, sx, sy, pt, exact, ns, eps, partial, st));
    size_t base = c->nseg;
    for (const auto& p : c->parts) {
        HIPCHK(klaunch::sweep_exact(10 + p->alleles, p->mask.p, p->row_words, p->rows.p, p->v_pad, p->table.p,
                                    p->seg.p, p->nseg, sx, sy, pt, exact, ns, eps, partial + base * NPART * ns, st));
        base += p->nseg;
    }
    return KESTREL_OK;
}

namespace {

// Claim order of a pair list in sample tiles: a stable counting sort by the
// tiles (min(x, y) / T, max(x, y) / T) of its samples, tile pairs in row-major
// order.  The pairs in flight at any time then come from a few tiles, so few
// sample rows are streamed per sweep (it pays where rows are likelihoods).
void tile_order(const uint32_t* pairs, size_t n, uint32_t n_samples, uint32_t T, std::vector<uint64_t>& order) {
    // at most 4,096 tiles a side (the bucket counters stay within 128 MB)
    T = std::max<uint32_t>(T, (n_samples + 4095) / 4096);
    const uint64_t nt = (n_samples + T - 1) / T;
    std::vector<uint64_t> start(nt * nt + 1, 0);
    auto key = [&](size_t i) {
        const uint32_t a = std::min(pairs[2 * i], pairs[2 * i + 1]) / T, b = std::max(pairs[2 * i], pairs[2 * i + 1]) / T;
        return (uint64_t)a * nt + b;
    };
    for (size_t i = 0; i < n; i++) start[key(i) + 1]++;
    for (uint64_t k = 0; k < nt * nt; k++) start[k + 1] += start[k];
    order.resize(n);
    for (size_t i = 0; i < n; i++) order[start[key(i)]++] = i;
}

// Pairs in flight per slot group that fill the device: 2 blocks of 256 lanes
// per CU over the cohort's segments.
uint64_t default_slots(const kestrel_cohort* c) {
    const uint64_t resident = (uint64_t)std::max(1, c->n_cu) * 2;
    return std::max<uint64_t>(1, resident / std::max<uint32_t>(1, c->nseg)) * SWEEP_TPB;
}

// ... of the class-sorted sweep: one workgroup of sixteen pairs (four
// wavefronts per SIMD) per CU, two device fills
uint64_t class_slots(const kestrel_cohort* c) {
    const uint64_t resident = (uint64_t)std::max(1, c->n_cu) * 16;
    return std::max<uint64_t>(1, resident / std::max<uint32_t>(1, c->nseg_cls)) * 2;
}

}  // namespace

int recondition_records(kestrel_cohort* c, const uint32_t* geno, uint32_t row_words, const uint32_t* seg,
                        uint32_t nseg, const kestrel_tune& tune, kestrel_output* recs, size_t n,
                        const std::function<void(size_t, uint32_t&, uint32_t&)>& rows, hipStream_t st,
                        uint64_t* n_done) {
    std::vector<size_t> idx;
    for (size_t i = 0; i < n; i++)
        if (recs[i].status == KESTREL_PAIR_OK && recs[i].iters < tune.sqp_max_iter &&
            !(recs[i].cond >= 0.0 && recs[i].cond < RECOND_MIN))
            idx.push_back(i);
    if (n_done) *n_done = idx.size();
    if (idx.empty()) return KESTREL_OK;
    // up to eight device fills of pairs per launch
    const size_t batch = std::min<size_t>(idx.size(), std::max<size_t>(8 * default_slots(c), SWEEP_TPB));
    const uint32_t ns = (uint32_t)((batch + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB);
    // the tri/tetra-allelic parts' segments follow the main part's
    uint32_t nall = nseg;
    for (const auto& p : c->parts) nall += p->nseg;
    DevBuf<uint32_t> dsx, dsy;
    DevBuf<int> dact;
    DevBuf<double> dpt, dpart, dcond;
    HIPCHK(dsx.alloc(ns));
    HIPCHK(dsy.alloc(ns));
    HIPCHK(dact.alloc(ns));
    HIPCHK(dpt.alloc((size_t)9 * ns));
    HIPCHK(dpart.alloc((size_t)nall * NPART * ns));
    HIPCHK(dcond.alloc(ns));
    std::vector<uint32_t> hx(ns), hy(ns);
    std::vector<int> hact(ns);
    std::vector<double> hpt((size_t)9 * ns), hcond(ns);
    for (size_t b0 = 0; b0 < idx.size(); b0 += batch) {
        const size_t m = std::min(batch, idx.size() - b0);
        std::fill(hx.begin(), hx.end(), 0u);
        std::fill(hy.begin(), hy.end(), 0u);
        std::fill(hact.begin(), hact.end(), 0);
        std::fill(hpt.begin(), hpt.end(), 0.0);
        for (size_t j = 0; j < m; j++) {
            const kestrel_output& o = recs[idx[b0 + j]];
            rows(idx[b0 + j], hx[j], hy[j]);
            hact[j] = 1;
            for (int k = 0; k < 9; k++) hpt[(size_t)k * ns + j] = o.jacquard[k];
        }
        HIPCHK(hipMemcpyAsync(dsx.p, hx.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dsy.p, hy.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dact.p, hact.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dpt.p, hpt.data(), hpt.size() * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(klaunch::sweep(geno, row_words, c->af.p, seg, nseg, dsx.p, dsy.p, dpt.p, dact.p, ns, tune.epsilon,
                              dpart.p, st, c->soft ? c->glrows.p : nullptr, c->v_pad, c->gl_stride, false, false));
        size_t off = nseg;
        for (const auto& p : c->parts) {  // FP64 sweeps of the tri/tetra-allelic parts
            HIPCHK(klaunch::sweep_multi(p->alleles, p->mask.p, p->row_words, p->rows.p, p->v_pad, p->table.p,
                                        p->af.p, p->seg.p, p->nseg, dsx.p, dsy.p, dpt.p, dact.p, ns, tune.epsilon,
                                        dpart.p + off * NPART * ns, st, nullptr, p->codes.n ? p->codes.p : nullptr, p->lut.p));
            off += p->nseg;
        }
        HIPCHK(klaunch::recondition(dpart.p, nall, ns, (uint32_t)m, dpt.p, tune.sqp_conv_tol, dcond.p, st));
        HIPCHK(hipMemcpyAsync(hcond.data(), dcond.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (size_t j = 0; j < m; j++) recs[idx[b0 + j]].cond = hcond[j];
    }
    return KESTREL_OK;
}

namespace {

int open_device(kestrel_cohort* c) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KESTREL_EDEVICE, "no HIP device: libkestrel_amd needs an MI355X (gfx950) GPU");
    HIPCHK(hipGetDevice(&c->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    c->n_cu = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    return KESTREL_OK;
}

}  // namespace

extern "C" {

const char* kestrel_last_error(void) { return g_err.c_str(); }
const char* kestrel_version(void) { return "kestrel_amd 0.1.0 (gfx950)"; }

int kestrel_tune_default(kestrel_tune* t, uint64_t iterations) {
    if (!t || iterations == 0) return fail(KESTREL_EINVAL, "iterations must be non-zero");
    t->sqp_max_iter = iterations;
    t->sqp_conv_tol = 1e-8;
    t->qp_max_iter = 10;
    t->qp_conv_tol = 1e-10;
    t->qp_zero_search_tol = 1e-12;
    t->bls_max_iter = 10;
    t->bls_sufficient_decrease = 1e-4;
    t->bls_step_size_reduce = 0.9;
    t->epsilon = 1e-8;
    return KESTREL_OK;
}

int kestrel_cohort_create(kestrel_cohort** out, const uint8_t* packed, uint32_t n_samples, uint32_t n_sites,
                          size_t row_stride_bytes, const double* af, double maf, uint64_t af_iterations,
                          int precision) {
    if (!out || !packed) return fail(KESTREL_EINVAL, "null argument");
    *out = nullptr;
    if (n_samples == 0 || n_sites == 0) return fail(KESTREL_EINVAL, "need at least one sample and one site");
    if (row_stride_bytes < (n_sites + 3) / 4) return fail(KESTREL_EINVAL, "row stride too small");
    if (precision != KESTREL_FP64 && precision != KESTREL_MIXED)
        return fail(KESTREL_EINVAL, "unknown precision %d", precision);
    if (maf == KESTREL_MAF_DEFAULT) maf = 1.0 / (2.0 * (double)n_samples);  // wrapper.rs:101
    if (!(maf >= 0.0 && maf <= 1.0)) return fail(KESTREL_EINVAL, "Minimum allele frequency must be within 0 to 1");
    if (!af && af_iterations == 0) return fail(KESTREL_EINVAL, "iterations must be non-zero");

    auto c = new kestrel_cohort();
    std::unique_ptr<kestrel_cohort> guard(c);
    int rc = open_device(c);
    if (rc) return rc;
    c->n_samples = n_samples;
    c->n_sites = n_sites;
    c->mixed = precision == KESTREL_MIXED;

    DevBuf<uint8_t> dpacked;
    HIPCHK(dpacked.alloc((size_t)n_samples * row_stride_bytes));
    HIPCHK(hipMemcpyAsync(dpacked.p, packed, (size_t)n_samples * row_stride_bytes, hipMemcpyHostToDevice,
                          c->stream));

    c->af_all.resize((size_t)n_sites * 2);
    if (af) {
        std::memcpy(c->af_all.data(), af, c->af_all.size() * sizeof(double));
    } else {
        DevBuf<double> daf;
        DevBuf<int> dstatus;
        HIPCHK(daf.alloc((size_t)n_sites * 2));
        HIPCHK(dstatus.alloc(1));
        HIPCHK(hipMemsetAsync(dstatus.p, 0, sizeof(int), c->stream));
        kestrel_tune t;
        kestrel_tune_default(&t, af_iterations);
        t.qp_max_iter = 3;  // allele/mod.rs:93: A + 1
        HIPCHK(klaunch::af(dpacked.p, row_stride_bytes, n_samples, n_sites, to_dev(t), maf, daf.p, dstatus.p,
                           c->stream));
        int status = 0;
        HIPCHK(hipMemcpyAsync(c->af_all.data(), daf.p, c->af_all.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(&status, dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (status) return fail(KESTREL_EDATA, "Must have at least one row to analyze (a site has no called sample)");
    }

    // wrapper.rs:352-410: keep sites with two non-zero allele frequencies
    c->kept_mask.assign(n_sites, 0);
    std::vector<uint32_t> kept;
    kept.reserve(n_sites);
    for (uint32_t s = 0; s < n_sites; s++)
        if (c->af_all[2 * s] > 0.0 && c->af_all[2 * s + 1] > 0.0) {
            c->kept_mask[s] = 1;
            kept.push_back(s);
        }
    rc = repack_kept(c, dpacked.p, row_stride_bytes, kept);
    if (rc) return rc;
    if (c->n_kept) {
        DevBuf<unsigned long long> dmiss;
        HIPCHK(dmiss.alloc(1));
        HIPCHK(hipMemsetAsync(dmiss.p, 0, sizeof(unsigned long long), c->stream));
        HIPCHK(klaunch::count_missing(c->geno.p, n_samples, c->row_words, c->n_kept, dmiss.p, c->stream));
        unsigned long long miss = 0;
        HIPCHK(hipMemcpyAsync(&miss, dmiss.p, sizeof(miss), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->missing_frac = (double)miss / ((double)n_samples * (double)c->n_kept);
    }
    std::vector<double> afk((size_t)c->v_pad * 2, 0.5);
    for (uint32_t t = 0; t < c->n_kept; t++) {
        afk[2 * t] = c->af_all[2 * kept[t]];
        afk[2 * t + 1] = c->af_all[2 * kept[t] + 1];
    }
    HIPCHK(c->af.alloc(afk.size()));
    HIPCHK(hipMemcpyAsync(c->af.p, afk.data(), afk.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = c->set_segments();
    if (rc) return rc;
    *out = guard.release();
    return KESTREL_OK;
}

int kestrel_cohort_create_gl(kestrel_cohort** out, const double* gl, uint32_t n_samples, uint32_t n_sites,
                             const double* af, double maf, uint64_t af_iterations) {
    if (!out || !gl) return fail(KESTREL_EINVAL, "null argument");
    *out = nullptr;
    if (n_samples == 0 || n_sites == 0) return fail(KESTREL_EINVAL, "need at least one sample and one site");
    if (maf == KESTREL_MAF_DEFAULT) maf = 1.0 / (2.0 * (double)n_samples);  // wrapper.rs:101
    if (!(maf >= 0.0 && maf <= 1.0)) return fail(KESTREL_EINVAL, "Minimum allele frequency must be within 0 to 1");
    if (!af && af_iterations == 0) return fail(KESTREL_EINVAL, "iterations must be non-zero");

    auto c = new kestrel_cohort();
    std::unique_ptr<kestrel_cohort> guard(c);
    int rc = open_device(c);
    if (rc) return rc;
    c->soft = true;
    c->n_samples = n_samples;
    c->n_sites = n_sites;
    const size_t n_vals = (size_t)n_sites * n_samples * 3;
    c->lmax = 0.0;
    rc = scan_likelihoods(gl, n_vals, &c->lmax);
    if (rc) return rc;
    DevBuf<double> dgl;
    HIPCHK(dgl.alloc(n_vals));
    HIPCHK(hipMemcpyAsync(dgl.p, gl, n_vals * sizeof(double), hipMemcpyHostToDevice, c->stream));

    c->af_all.resize((size_t)n_sites * 2);
    if (af) {
        std::memcpy(c->af_all.data(), af, c->af_all.size() * sizeof(double));
    } else {
        DevBuf<double> daf;
        DevBuf<int> dstatus;
        HIPCHK(daf.alloc((size_t)n_sites * 2));
        HIPCHK(dstatus.alloc(1));
        HIPCHK(hipMemsetAsync(dstatus.p, 0, sizeof(int), c->stream));
        kestrel_tune t;
        kestrel_tune_default(&t, af_iterations);
        t.qp_max_iter = 3;  // allele/mod.rs:93: A + 1
        HIPCHK(klaunch::af_gl(dgl.p, n_samples, n_sites, to_dev(t), maf, daf.p, dstatus.p, c->stream));
        int status = 0;
        HIPCHK(hipMemcpyAsync(c->af_all.data(), daf.p, c->af_all.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(&status, dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (status) return fail(KESTREL_EDATA, "Must have at least one row to analyze (a site has no called sample)");
    }

    // wrapper.rs:352-410: keep sites with two non-zero allele frequencies
    c->kept_mask.assign(n_sites, 0);
    std::vector<uint32_t> kept;
    for (uint32_t s = 0; s < n_sites; s++)
        if (c->af_all[2 * s] > 0.0 && c->af_all[2 * s + 1] > 0.0) {
            c->kept_mask[s] = 1;
            kept.push_back(s);
        }
    c->n_kept = (uint32_t)kept.size();
    c->v_pad = std::max<uint32_t>(SITE_PAD, (c->n_kept + SITE_PAD - 1) / SITE_PAD * SITE_PAD);
    c->row_words = c->v_pad / 16;
    DevBuf<uint32_t> dkept;
    HIPCHK(dkept.alloc(std::max<size_t>(kept.size(), 1)));
    if (!kept.empty())
        HIPCHK(hipMemcpyAsync(dkept.p, kept.data(), kept.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              c->stream));
    HIPCHK(c->geno.alloc((size_t)n_samples * c->row_words));
    // 16-byte rows when every called row holds an exact 1.0 (likelihoods
    // scaled by their maximum, as vcf.rs:349-368 makes GL rows and a PL of 0
    // makes a PL row), else rows of three likelihoods
    DevBuf<int> no_one;
    HIPCHK(no_one.alloc(1));
    HIPCHK(hipMemsetAsync(no_one.p, 0, sizeof(int), c->stream));
    HIPCHK(c->glrows.alloc((size_t)n_samples * c->v_pad * 2));
    HIPCHK(klaunch::gather_gl(dgl.p, n_samples, dkept.p, c->n_kept, c->glrows.p, c->v_pad, c->geno.p, c->row_words,
                              1, no_one.p, c->stream));
    int lacks = 0;
    HIPCHK(hipMemcpyAsync(&lacks, no_one.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->gl_stride = 2;
    if (lacks) {
        HIPCHK(c->glrows.alloc((size_t)n_samples * c->v_pad * 3));
        HIPCHK(klaunch::gather_gl(dgl.p, n_samples, dkept.p, c->n_kept, c->glrows.p, c->v_pad, c->geno.p,
                                  c->row_words, 0, no_one.p, c->stream));
        c->gl_stride = 3;
    }
    std::vector<double> afk((size_t)c->v_pad * 2, 0.5);
    for (uint32_t t = 0; t < c->n_kept; t++) {
        afk[2 * t] = c->af_all[2 * kept[t]];
        afk[2 * t + 1] = c->af_all[2 * kept[t] + 1];
    }
    HIPCHK(c->af.alloc(afk.size()));
    HIPCHK(hipMemcpyAsync(c->af.p, afk.data(), afk.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = codes_if_one_hot(c);
    if (rc) return rc;
    rc = c->set_segments();
    if (rc) return rc;
    *out = guard.release();
    return KESTREL_OK;
}

int kestrel_cohort_create_gl_multi(kestrel_cohort** out, uint32_t n_samples, const double* gl2, uint32_t n2,
                                   const double* gl3, uint32_t n3, const double* gl4, uint32_t n4, double maf,
                                   uint64_t af_iterations) {
    if (!out) return fail(KESTREL_EINVAL, "null argument");
    *out = nullptr;
    const uint32_t nin[3] = {n2, n3, n4};
    const double* gin[3] = {gl2, gl3, gl4};
    if (n_samples == 0 || (uint64_t)n2 + n3 + n4 == 0)
        return fail(KESTREL_EINVAL, "need at least one sample and one site");
    for (int b = 0; b < 3; b++)
        if (nin[b] && !gin[b]) return fail(KESTREL_EINVAL, "null likelihoods for %u sites", nin[b]);
    if (maf == KESTREL_MAF_DEFAULT) maf = 1.0 / (2.0 * (double)n_samples);  // wrapper.rs:101
    if (!(maf >= 0.0 && maf <= 1.0)) return fail(KESTREL_EINVAL, "Minimum allele frequency must be within 0 to 1");
    if (af_iterations == 0) return fail(KESTREL_EINVAL, "iterations must be non-zero");

    auto c = new kestrel_cohort();
    std::unique_ptr<kestrel_cohort> guard(c);
    int rc = open_device(c);
    if (rc) return rc;
    c->soft = true;
    c->multi = true;
    c->n_samples = n_samples;
    c->n_sites = n2 + n3 + n4;
    c->lmax = 0.0;
    for (int b = 0; b < 3; b++) {
        rc = scan_likelihoods(gin[b], (size_t)nin[b] * n_samples * ((b + 2) * (b + 3) / 2), &c->lmax);
        if (rc) return rc;
    }

    // allele frequencies per input bucket (allele/mod.rs:37-134)
    DevBuf<double> dgl[3];
    DevBuf<int> dstatus;
    HIPCHK(dstatus.alloc(1));
    HIPCHK(hipMemsetAsync(dstatus.p, 0, sizeof(int), c->stream));
    for (int b = 0; b < 3; b++) {
        const int A = b + 2, G = A * (A + 1) / 2;
        const size_t vals = (size_t)nin[b] * n_samples * G;
        HIPCHK(dgl[b].alloc(vals));
        c->af_in[b].assign((size_t)nin[b] * A, 0.0);
        if (!nin[b]) continue;
        HIPCHK(hipMemcpyAsync(dgl[b].p, gin[b], vals * sizeof(double), hipMemcpyHostToDevice, c->stream));
        DevBuf<double> daf;
        HIPCHK(daf.alloc((size_t)nin[b] * A));
        kestrel_tune t;
        kestrel_tune_default(&t, af_iterations);
        t.qp_max_iter = (uint64_t)A + 1;  // allele/mod.rs:93
        HIPCHK(klaunch::af_gl(dgl[b].p, n_samples, nin[b], to_dev(t), maf, daf.p, dstatus.p, c->stream, A));
        HIPCHK(hipMemcpyAsync(c->af_in[b].data(), daf.p, c->af_in[b].size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    int status = 0;
    HIPCHK(hipMemcpy(&status, dstatus.p, sizeof(int), hipMemcpyDeviceToHost));
    if (status) return fail(KESTREL_EDATA, "Must have at least one row to analyze (a site has no called sample)");

    // wrapper.rs:237-410 remove_zero_frequencies: drop alleles of frequency 0,
    // re-bucket every site by its remaining allele count (buckets of 2, 3, 4
    // alleles in turn, sites in order; one allele left drops the site)
    struct PartMap {
        std::vector<uint32_t> site;
        std::vector<uint8_t> alleles, gmap;
        std::vector<double> af;
    } pm[3];
    for (int b = 0; b < 3; b++) {
        const int A = b + 2;
        for (uint32_t s = 0; s < nin[b]; s++) {
            int nz[4], m = 0;
            for (int a = 0; a < A; a++)
                if (c->af_in[b][(size_t)A * s + a] > 0.0) nz[m++] = a;
            if (m < 2) continue;
            PartMap& p = pm[m - 2];
            p.site.push_back(s);
            p.alleles.push_back((uint8_t)A);
            for (int jj = 0; jj < m; jj++)
                for (int ii = 0; ii <= jj; ii++) p.gmap.push_back((uint8_t)(nz[jj] * (nz[jj] + 1) / 2 + nz[ii]));
            for (int q = 0; q < m; q++) p.af.push_back(c->af_in[b][(size_t)A * s + nz[q]]);
        }
    }
    for (int q = 0; q < 3; q++) c->part_sites[q] = (uint32_t)pm[q].site.size();

    // compact: 16-byte rows of the 2-allele part (as create_gl makes them);
    // *lacks reports a called row with no exact 1.0
    auto build_part = [&](int q, uint32_t v_pad, DevBuf<double>& rows, DevBuf<uint32_t>& mask, DevBuf<double>& af,
                          int compact, int* lacks) -> int {
        const int A = q + 2, G = A * (A + 1) / 2;
        const uint32_t n_t = (uint32_t)pm[q].site.size(), row_words = v_pad / 16;
        DevBuf<uint32_t> dsite;
        DevBuf<uint8_t> dal, dgm;
        HIPCHK(dsite.alloc(std::max<uint32_t>(1, n_t)));
        HIPCHK(dal.alloc(std::max<uint32_t>(1, n_t)));
        HIPCHK(dgm.alloc(std::max<size_t>(1, pm[q].gmap.size())));
        if (n_t) {
            HIPCHK(hipMemcpyAsync(dsite.p, pm[q].site.data(), n_t * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(dal.p, pm[q].alleles.data(), n_t, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(dgm.p, pm[q].gmap.data(), pm[q].gmap.size(), hipMemcpyHostToDevice, c->stream));
        }
        DevBuf<int> no_one;
        HIPCHK(no_one.alloc(1));
        HIPCHK(hipMemsetAsync(no_one.p, 0, sizeof(int), c->stream));
        HIPCHK(rows.alloc((size_t)n_samples * v_pad * (compact ? 2 : G)));
        HIPCHK(mask.alloc((size_t)n_samples * row_words));
        HIPCHK(klaunch::gather_multi(dgl[0].p, dgl[1].p, dgl[2].p, n_samples, dsite.p, dal.p, dgm.p, n_t, G, rows.p,
                                     v_pad, mask.p, row_words, c->stream, compact, no_one.p));
        if (lacks) HIPCHK(hipMemcpyAsync(lacks, no_one.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        std::vector<double> afp((size_t)v_pad * A, 1.0 / (double)A);  // padding sites are missing for every sample
        std::copy(pm[q].af.begin(), pm[q].af.end(), afp.begin());
        HIPCHK(af.alloc(afp.size()));
        HIPCHK(hipMemcpyAsync(af.p, afp.data(), afp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));  // the maps and host vectors go out of scope
        return KESTREL_OK;
    };
    // the 2-allele part is the main part (sweep_gl); 3 and 4 alleles follow
    c->n_kept = c->part_sites[0];
    c->v_pad = std::max<uint32_t>(SITE_PAD, (c->n_kept + SITE_PAD - 1) / SITE_PAD * SITE_PAD);
    c->row_words = c->v_pad / 16;
    int lacks = 0;
    rc = build_part(0, c->v_pad, c->glrows, c->geno, c->af, 1, &lacks);
    if (rc) return rc;
    c->gl_stride = 2;
    if (lacks) {
        rc = build_part(0, c->v_pad, c->glrows, c->geno, c->af, 0, nullptr);
        if (rc) return rc;
        c->gl_stride = 3;
    }
    rc = codes_if_one_hot(c);  // the 2-allele part (its rows are the main part's)
    if (rc) return rc;
    for (int q = 1; q < 3; q++) {
        if (!c->part_sites[q]) continue;
        std::unique_ptr<SitePart> p(new SitePart());
        p->alleles = q + 2;
        p->n_sites = c->part_sites[q];
        p->v_pad = (p->n_sites + SITE_PAD - 1) / SITE_PAD * SITE_PAD;
        p->row_words = p->v_pad / 16;
        rc = build_part(q, p->v_pad, p->rows, p->mask, p->af, 0, nullptr);
        if (rc) return rc;
        HIPCHK(p->table.alloc((size_t)p->v_pad * klaunch::iis_table_size(p->alleles)));
        HIPCHK(klaunch::iis_table(p->alleles, p->af.p, p->v_pad, p->table.p, c->stream));
        HIPCHK(p->table32.alloc(p->table.n));
        HIPCHK(klaunch::round_f32(p->table.p, p->table32.p, p->table.n, c->stream));
        // hard calls fed as likelihoods: the FP64 sweeps read each genotype
        // pair's coefficients by the 4-bit indices (the rows stay for the
        // mixed form and evaluations outside the sweeps' domain)
        DevBuf<int> bad;
        HIPCHK(bad.alloc(1));
        HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
        HIPCHK(p->codes.alloc((size_t)c->n_samples * (p->v_pad / 8)));
        HIPCHK(klaunch::part_codes(p->alleles, p->rows.p, p->mask.p, p->row_words, p->v_pad, c->n_samples,
                                   p->codes.p, bad.p, c->stream));
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (hbad) {
            p->codes.release();
        } else {
            HIPCHK(p->lut.alloc((size_t)p->v_pad * klaunch::part_lut_size(p->alleles)));
            HIPCHK(klaunch::part_lut(p->alleles, p->table.p, p->v_pad, p->lut.p, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->parts.push_back(std::move(p));
    }
    rc = c->set_segments();
    if (rc) return rc;
    *out = guard.release();
    return KESTREL_OK;
}

int kestrel_cohort_multi_info(const kestrel_cohort* c, uint32_t part_sites[3], double* af2, double* af3,
                              double* af4) {
    if (!c) return fail(KESTREL_EINVAL, "null cohort");
    if (!c->multi) return fail(KESTREL_EINVAL, "cohort was not created by kestrel_cohort_create_gl_multi or _gt_multi");
    if (part_sites)
        for (int q = 0; q < 3; q++) part_sites[q] = c->part_sites[q];
    double* dst[3] = {af2, af3, af4};
    for (int b = 0; b < 3; b++)
        if (dst[b]) std::copy(c->af_in[b].begin(), c->af_in[b].end(), dst[b]);
    return KESTREL_OK;
}

int kestrel_cohort_allele_freqs(const kestrel_cohort* c, double* af_out, uint8_t* kept_out) {
    if (!c) return fail(KESTREL_EINVAL, "null cohort");
    if (c->multi) return fail(KESTREL_EINVAL, "a multi-allelic cohort reports frequencies by kestrel_cohort_multi_info");
    if (af_out) std::memcpy(af_out, c->af_all.data(), c->af_all.size() * sizeof(double));
    if (kept_out) std::memcpy(kept_out, c->kept_mask.data(), c->kept_mask.size());
    return KESTREL_OK;
}

int kestrel_cohort_n_kept(const kestrel_cohort* c, uint32_t* n_kept) {
    if (!c || !n_kept) return fail(KESTREL_EINVAL, "null argument");
    *n_kept = c->n_kept;
    return KESTREL_OK;
}

int kestrel_set_option(kestrel_cohort* c, const char* key, int64_t value) {
    if (!c || !key) return fail(KESTREL_EINVAL, "null argument");
    if (!strcmp(key, "slots")) {
        if (value < 0) return fail(KESTREL_EINVAL, "slots must be >= 0");
        c->opt_slots = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "segment_sites")) {
        if (value < CHUNK || value % CHUNK) return fail(KESTREL_EINVAL, "segment_sites must be a multiple of %d", CHUNK);
        c->opt_segment_sites = value;
        return c->set_segments();
    }
    if (!strcmp(key, "streams")) {
        if (value != 1 && value != 2) return fail(KESTREL_EINVAL, "streams must be 1 or 2");
        c->opt_streams = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "pair_tile")) {
        if (value < -1 || value > (int64_t)1 << 30) return fail(KESTREL_EINVAL, "pair_tile must be -1, 0 or a tile size");
        c->opt_pair_tile = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "precision")) {
        if (value != KESTREL_FP64 && value != KESTREL_MIXED)
            return fail(KESTREL_EINVAL, "precision must be KESTREL_FP64 or KESTREL_MIXED");
        if (value == KESTREL_MIXED && c->gt)
            return fail(KESTREL_EINVAL, "the mixed-precision form covers the likelihood fit, not -f GT cohorts");
        c->mixed = value == KESTREL_MIXED;
        return KESTREL_OK;
    }
    if (!strcmp(key, "pair_lanes")) {
        if (value != 1 && value != 64 && value != -1) return fail(KESTREL_EINVAL, "pair_lanes must be 1, 64 or -1");
        c->opt_pair_lanes = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "class_sweep")) {
        if (value != 0 && value != 1) return fail(KESTREL_EINVAL, "class_sweep must be 0 or 1");
        c->opt_class_sweep = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "class_waves")) {
        if (value != 8 && value != 12 && value != 14 && value != 16)
            return fail(KESTREL_EINVAL, "class_waves must be 8, 12, 14 or 16");
        c->opt_class_waves = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "class_list_mib")) {
        if (value < 0) return fail(KESTREL_EINVAL, "class_list_mib must be >= 0");
        c->opt_class_list_mib = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "skip_masked")) {
        if (value < -1 || value > 1) return fail(KESTREL_EINVAL, "skip_masked must be -1, 0 or 1");
        c->opt_skip_masked = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "graph")) {
        if (value != 0 && value != 1) return fail(KESTREL_EINVAL, "graph must be 0 or 1");
        c->opt_graph = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "batch")) {
        if (value < 1 || value > 64) return fail(KESTREL_EINVAL, "batch must be within 1 to 64");
        c->opt_batch = value;
        return KESTREL_OK;
    }
    if (!strcmp(key, "serve_policy")) {
        if (value != 0 && value != 1) return fail(KESTREL_EINVAL, "serve_policy must be 0 or 1");
        c->opt_serve_policy = value;
        return KESTREL_OK;
    }
    return fail(KESTREL_EINVAL, "unknown option %s", key);
}

int kestrel_last_fit_stats(const kestrel_cohort* c, double* sweep_ms, uint64_t* sweeps, double* wall_ms) {
    if (!c) return fail(KESTREL_EINVAL, "null cohort");
    if (sweep_ms) *sweep_ms = c->last_sweep_ms;
    if (sweeps) *sweeps = c->last_sweeps;
    if (wall_ms) *wall_ms = c->last_wall_ms;
    return KESTREL_OK;
}

int kestrel_last_fit_busy(const kestrel_cohort* c, double* busy_ms, uint32_t* groups) {
    if (!c) return fail(KESTREL_EINVAL, "null cohort");
    if (busy_ms) *busy_ms = c->last_busy_ms;
    if (groups) *groups = (uint32_t)c->last_groups;
    return KESTREL_OK;
}

int kestrel_last_fit_recondition(const kestrel_cohort* c, double* ms, uint64_t* pairs) {
    if (!c) return fail(KESTREL_EINVAL, "null cohort");
    if (ms) *ms = c->last_recond_ms;
    if (pairs) *pairs = c->last_recond_pairs;
    return KESTREL_OK;
}

int kestrel_last_fit_form(const kestrel_cohort* c, int* form) {
    if (!c || !form) return fail(KESTREL_EINVAL, "null argument");
    *form = c->last_form;
    return KESTREL_OK;
}

int kestrel_fit_pairs(kestrel_cohort* c, const kestrel_tune* tune, const uint32_t* pairs, size_t n_pairs,
                      uint32_t n_select, kestrel_output* out) {
    if (!c || !tune || !out) return fail(KESTREL_EINVAL, "null argument");
    if (c->gt) return fail(KESTREL_EINVAL, "a GT cohort is fitted with kestrel_fit_pairs_gt");
    if (tune->sqp_max_iter == 0 || tune->qp_max_iter == 0 || tune->bls_max_iter == 0)
        return fail(KESTREL_EINVAL, "iteration limits must be non-zero");
    // the fit's points lie on the simplex: with eps in [1e-18, 1e14] every
    // L = eps + c.D is a positive normal number, as are 1/L and the product of
    // 16 of them, the sweeps' domain
    if (!eps_in_domain(tune->epsilon) || c->lmax * c->lmax > 1e14)
        return fail(KESTREL_EINVAL, "epsilon %g outside [1e-18, 1e14] (or likelihoods above 1e7) for a pair fit",
                    tune->epsilon);
    std::vector<uint32_t> tri;
    if (!pairs) {
        if (n_select > c->n_samples)  // coefficients.rs:53-55
            return fail(KESTREL_EINVAL, "Requested samples %u exceeds total samples %u", n_select, c->n_samples);
        n_pairs = (size_t)n_select * ((size_t)n_select + 1) / 2;
        tri.resize(2 * n_pairs);
        size_t i = 0;
        for (uint32_t x = 0; x < n_select; x++)
            for (uint32_t y = x; y < n_select; y++, i++) {
                tri[2 * i] = x;
                tri[2 * i + 1] = y;
            }
        pairs = tri.data();
    } else {
        for (size_t i = 0; i < 2 * n_pairs; i++)
            if (pairs[i] >= c->n_samples) return fail(KESTREL_EINVAL, "pair %zu names sample %u of %u", i / 2, pairs[i],
                                                      c->n_samples);
    }
    if (n_pairs == 0) return KESTREL_OK;
    uint64_t usable_sites = c->n_kept;
    for (const auto& p : c->parts) usable_sites += p->n_sites;
    if (usable_sites == 0) {
        const double nan = std::nan("");
        for (size_t i = 0; i < n_pairs; i++) {
            kestrel_output& o = out[i];
            o.x = pairs[2 * i];
            o.y = pairs[2 * i + 1];
            for (double& v : o.jacquard) v = nan;
            o.obj = nan;
            o.iters = 0;
            o.cond = nan;
            o.n_sites = 0;
            o.status = KESTREL_PAIR_NO_SITES;
        }
        return KESTREL_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();

    // claims in runs and sweeps that pass over the sites where a wavefront's
    // shared sample is missing (2-bit cohorts; option "skip_masked")
    // (automatic for the FP64 form only: the mixed form's pass holds two
    // wavefronts per SIMD instead of three, and measured 1,241,070 against
    // 1,266,335 pair-fits/s without it on C3)
    // a pair's sites across a wavefront (option "pair_lanes" 64): FP64 fits of
    // 2-bit and biallelic likelihood-row cohorts; claims in list order (the
    // pass has no shared sample to skip on)
    // (-1: automatic, for fits of at most LANES_AUTO_PAIRS pairs: measured on
    // 1,000 x 100,000 sites, 100 / 999 / 3,996 / 31,968 pairs fit in 1.9 /
    // 8.4 / 24.6 / 162 ms against 18.5 / 21.5 / 38.0 / 110 ms one pair a lane)
    const bool lanes_form = c->parts.empty() && !c->mixed;
    if (c->opt_pair_lanes == 64 && !lanes_form)
        return fail(KESTREL_EINVAL, "pair_lanes 64 covers FP64 fits of 2-bit and biallelic likelihood-row cohorts");
    const bool lanes = c->opt_pair_lanes == 64 || (c->opt_pair_lanes < 0 && lanes_form && n_pairs <= LANES_AUTO_PAIRS);
    // the class-sorted sweep (option "class_sweep"): one pair a wavefront over
    // its sites sorted by genotype-pair pattern; the sweep passes over missing
    // calls by construction, so claims in runs add nothing
    // its lists take ~2 bytes a site a slot: the slots are capped by the
    // device memory
    const bool use_class = !lanes && c->class_form();
    const uint64_t class_cap = use_class ? class_slot_cap(c->class_blocks(), 2, c->opt_class_list_mib) : 0;
    if (use_class && class_cap == 0)
        return fail(KESTREL_EDEVICE, "device memory for the class-sorted sweep's lists of 2 x 256 slots (%llu bytes) "
                    "is not free (option class_sweep 0 selects the site-table sweep)",
                    (unsigned long long)(512 * ClassLists::slot_bytes(c->class_blocks())));
    const bool use_runs = !lanes && !use_class && !c->soft && c->parts.empty() &&
                          (c->opt_skip_masked == 1 ||
                           (c->opt_skip_masked < 0 && !c->mixed && c->missing_frac >= 0.02));
    // pairs in flight per group: fill the device unless set (twice with runs:
    // workgroups then end at different times, and a launch of one round would
    // last as long as its slowest; three times for the 2-bit mixed form
    // without runs, whose sweep fits three wavefronts per SIMD: measured
    // 571,872 -> 593,813 on C2 and 596,779 -> 657,444 on C4; its pass over
    // masked sites fits two), never more than the pairs available
    const uint64_t fills = use_runs ? 2 : (c->mixed && !c->soft) ? 3 : 1;
    uint64_t nslots = c->opt_slots > 0 ? (uint64_t)c->opt_slots
                      : use_class        ? class_slots(c)
                                         : default_slots(c) * fills;
    nslots = std::min<uint64_t>(nslots, ((uint64_t)n_pairs + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB);
    nslots = (nslots + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB;
    if (use_class) nslots = std::min(nslots, class_cap);
    const uint32_t ns = (uint32_t)nslots;
    // Two slot groups on two streams when there is work for both: while one
    // group folds and steps, the other's sweep keeps the device busy.  Both
    // claim from one pair counter; a pair's result does not depend on the
    // group or slot that fits it.
    const int ngroups = (c->opt_streams >= 2 && (uint64_t)n_pairs >= 4 * nslots) ? 2 : 1;

    DevBuf<uint32_t> dpairs;
    DevBuf<int> dlive;
    DevBuf<unsigned long long> dnext;
    DevBuf<kestrel_output> dout;
    HIPCHK(dpairs.alloc(2 * n_pairs));
    HIPCHK(dlive.alloc(1));
    HIPCHK(dnext.alloc(1));
    HIPCHK(dout.alloc(n_pairs));
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(dpairs.p, pairs, 2 * n_pairs * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    // claim order (option "pair_tile"; automatic where the sweep streams
    // likelihood rows: tiles of sqrt(pairs in flight) samples, so that the
    // pairs in flight span about one tile pair (measured on C2 soft: 16 ->
    // 245k, 64 -> 273k, 128 -> 276k, 256 -> 263k pair fits/s, list order 218k);
    // list order for 2-bit rows, which stay in cache)
    const int64_t auto_tile = std::min<int64_t>(
        256, std::max<int64_t>(32, (int64_t)std::lround(std::sqrt((double)ns * ngroups))));
    const int64_t tile =
        c->opt_pair_tile >= 0 ? c->opt_pair_tile : (c->soft || !c->parts.empty() ? auto_tile : 0);
    std::vector<uint64_t> order;
    DevBuf<uint64_t> dorder;
    const uint64_t* dorder_p = nullptr;
    if (tile > 0 && n_pairs > ns) {
        tile_order(pairs, n_pairs, c->n_samples, (uint32_t)tile, order);
        HIPCHK(dorder.alloc(n_pairs));
        HIPCHK(hipMemcpyAsync(dorder.p, order.data(), n_pairs * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        dorder_p = dorder.p;
    }
    DevBuf<uint64_t> druns;
    RunClaim rclaim;
    if (use_runs) {
        // runs of the claim order, at least one workgroup (256 claims) long,
        // cut where the first sample changes
        std::vector<uint64_t> runs;
        auto first_sample = [&](uint64_t idx) { return pairs[2 * (order.empty() ? idx : order[idx])]; };
        uint64_t begin = 0;
        for (uint64_t idx = 1; idx < n_pairs; idx++)
            if (idx - begin >= SWEEP_TPB && first_sample(idx) != first_sample(idx - 1)) {
                runs.push_back(begin);
                runs.push_back(idx);
                begin = idx;
            }
        runs.push_back(begin);
        runs.push_back(n_pairs);
        HIPCHK(druns.alloc(runs.size()));
        HIPCHK(hipMemcpyAsync(druns.p, runs.data(), runs.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));  // runs goes out of scope
        rclaim.runs = druns.p;
        rclaim.nruns = runs.size() / 2;
    }
    HIPCHK(hipMemsetAsync(dnext.p, 0, sizeof(unsigned long long), st));
    const int live0 = (int)ns * ngroups;
    HIPCHK(hipMemcpyAsync(dlive.p, &live0, sizeof(int), hipMemcpyHostToDevice, st));
    kdev::Tune dt = to_dev(*tune);
    if (c->mixed) dt.recond_margin = RECOND_MARGIN;
    const uint32_t nseg_fit = use_class ? c->nseg_class_total() : c->nseg_total();

    // sweeps and steps are queued BATCH at a time between host synchronisations;
    // each step adds the slots still live after it to hist[i], so the host
    // counts (and times) only sweeps that ran with live slots
    const int BATCH = (int)c->opt_batch;
    struct Group {
        hipStream_t stream = nullptr;
        DevBuf<uint32_t> sx, sy;
        DevBuf<double> pt, part;
        DevBuf<int> active, hist;
        DevBuf<SlotState> st;
        DevBuf<unsigned long long> wrun;  // claims in runs: each workgroup's unclaimed run part
        RunClaim rc;
        ClassLists cls;                   // the class-sorted sweep's pattern lists
        std::vector<hipEvent_t> ev;       // per batch buffer: start and end of each sweep
        std::vector<int> hist_h;
        hipEvent_t fin[2] = {nullptr, nullptr};  // per batch buffer: its live counts are on the host
        bool queued[2] = {false, false};          // per batch buffer: this group has a batch in it
        int live_before = 0;
        hipGraph_t graph = nullptr;     // option "graph": the BATCH sweeps and steps, captured once
        hipGraphExec_t exec = nullptr;
        ~Group() {
            for (hipEvent_t x : ev)
                if (x) (void)hipEventDestroy(x);
            for (hipEvent_t x : fin)
                if (x) (void)hipEventDestroy(x);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
        }
    };
    Group grp[2];
    hipEvent_t ref = nullptr;
    struct RefGuard {
        hipEvent_t& e;
        ~RefGuard() {
            if (e) (void)hipEventDestroy(e);
        }
    } refg{ref};
    HIPCHK(hipEventCreate(&ref));
    HIPCHK(hipStreamSynchronize(st));  // pair list and counters are set before either stream starts
    for (int g = 0; g < ngroups; g++) {
        Group& G = grp[g];
        G.stream = g == 0 ? c->stream : c->stream2;
        HIPCHK(G.sx.alloc(ns));
        HIPCHK(G.sy.alloc(ns));
        HIPCHK(G.pt.alloc((size_t)9 * ns));
        HIPCHK(G.part.alloc((size_t)nseg_fit * NPART * ns));
        if (use_class) {
            const int rc = G.cls.alloc(ns, c->class_blocks(), G.stream);
            if (rc) return rc;
        }
        HIPCHK(G.active.alloc(ns));
        HIPCHK(G.hist.alloc(2 * BATCH));
        HIPCHK(G.st.alloc(ns));
        G.ev.assign(4 * BATCH, nullptr);
        G.hist_h.assign(2 * BATCH, 0);
        for (hipEvent_t& x : G.ev) HIPCHK(hipEventCreate(&x));
        for (hipEvent_t& x : G.fin) HIPCHK(hipEventCreate(&x));
        HIPCHK(hipMemsetAsync(G.st.p, 0, (size_t)ns * sizeof(SlotState), G.stream));  // PH_IDLE
        HIPCHK(hipMemsetAsync(G.sx.p, 0, (size_t)ns * sizeof(uint32_t), G.stream));
        HIPCHK(hipMemsetAsync(G.sy.p, 0, (size_t)ns * sizeof(uint32_t), G.stream));
        HIPCHK(hipMemsetAsync(G.active.p, 0, (size_t)ns * sizeof(int), G.stream));
        if (use_runs) {
            HIPCHK(G.wrun.alloc((size_t)2 * (ns / SWEEP_TPB)));
            HIPCHK(hipMemsetAsync(G.wrun.p, 0, (size_t)2 * (ns / SWEEP_TPB) * sizeof(unsigned long long), G.stream));
            G.rc = rclaim;
            G.rc.wrun = G.wrun.p;
        }
        G.live_before = (int)ns;
    }
    HIPCHK(hipEventRecord(ref, st));
    if (ngroups == 2) HIPCHK(hipStreamWaitEvent(grp[1].stream, ref, 0));
    // claim the first pairs
    for (int g = 0; g < ngroups; g++) {
        Group& G = grp[g];
        HIPCHK(klaunch::step(G.st.p, G.sx.p, G.sy.p, G.pt.p, G.active.p, ns, G.part.p, nseg_fit, dpairs.p,
                             n_pairs, dorder_p, dnext.p, dout.p, dt, dlive.p, PairQueue{}, nullptr, G.stream, G.rc));
    }
    // option "graph": each group's BATCH sweeps and steps are captured once
    // into a HIP graph (kernel nodes only; the histogram is cleared outside
    // it) and relaunched per synchronisation; timing is then per batch
    const bool use_graph = c->opt_graph != 0;
    if (use_graph) {
        for (int g = 0; g < ngroups; g++) {
            Group& G = grp[g];
            HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
            int rs = KESTREL_OK;
            for (int i = 0; i < BATCH && rs == KESTREL_OK; i++) {
                rs = launch_sweeps(c, G.sx.p, G.sy.p, G.pt.p, G.active.p, ns, tune->epsilon, G.part.p, G.stream,
                                   use_runs, lanes, use_class ? &G.cls : nullptr);
                if (rs == KESTREL_OK) {
                    const hipError_t e = klaunch::step(G.st.p, G.sx.p, G.sy.p, G.pt.p, G.active.p, ns, G.part.p,
                                                       nseg_fit, dpairs.p, n_pairs, dorder_p, dnext.p,
                                                       dout.p, dt,
                                                       dlive.p, PairQueue{}, G.hist.p + i, G.stream, G.rc);
                    if (e != hipSuccess) rs = fail(KESTREL_EDEVICE, "HIP error %s", hipGetErrorString(e));
                }
            }
            const hipError_t ec = hipStreamEndCapture(G.stream, &G.graph);
            if (rs) return rs;
            HIPCHK(ec);
            HIPCHK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
        }
    }
    double sweep_ms = 0, exhaust_ms = INFINITY, end_ms = 0;
    uint64_t sweeps = 0, launched = 0;
    std::vector<std::pair<float, float>> spans;  // [start, end] of live sweeps, ms after ref
    // every pair needs at most 1 + sqp_max_iter * bls_max_iter evaluations
    const uint64_t per_pair = 1 + tune->sqp_max_iter * tune->bls_max_iter;
    const uint64_t max_sweeps = per_pair * (((uint64_t)n_pairs + ns - 1) / ns + 1) + 16;
    auto any_live = [&] {
        for (int g = 0; g < ngroups; g++)
            if (grp[g].live_before > 0) return true;
        return false;
    };
    // Batches are double-buffered (hist, events): without a graph, batch k + 1
    // is queued before the host waits for batch k, so the device does not drain
    // at each host synchronisation.  A group is queued while its last known live
    // count is positive; a batch queued after its group ran out sweeps and steps
    // no live slot (nothing is claimed or written) and is not counted.
    const bool pipe = !use_graph;
    auto enqueue = [&](int p) -> int {
        if (launched > max_sweeps) return fail(KESTREL_EDEVICE, "pair fit did not terminate after %llu sweeps",
                                               (unsigned long long)launched);
        for (int g = 0; g < ngroups; g++) {
            Group& G = grp[g];
            G.queued[p] = G.live_before > 0;
            if (G.queued[p]) HIPCHK(hipMemsetAsync(G.hist.p + p * BATCH, 0, BATCH * sizeof(int), G.stream));
        }
        for (int g = 0; g < ngroups && use_graph; g++) {
            Group& G = grp[g];
            if (!G.queued[p]) continue;
            HIPCHK(hipEventRecord(G.ev[p * 2 * BATCH], G.stream));
            HIPCHK(hipGraphLaunch(G.exec, G.stream));
            HIPCHK(hipEventRecord(G.ev[p * 2 * BATCH + 1], G.stream));
        }
        for (int i = 0; i < BATCH && !use_graph; i++)
            for (int g = 0; g < ngroups; g++) {
                Group& G = grp[g];
                if (!G.queued[p]) continue;
                // (the class lists of newly claimed pairs are built outside the
                // timed interval: it times the sweep kernel alone)
                if (use_class)
                    HIPCHK(klaunch::class_build(c->geno.p, c->row_words, G.sx.p, G.sy.p, G.active.p, ns, 0, ns,
                                                G.cls.nblk, G.cls.lists.p, G.cls.counts.p, G.cls.built.p, G.stream));
                HIPCHK(hipEventRecord(G.ev[p * 2 * BATCH + 2 * i], G.stream));
                int rs = launch_sweeps(c, G.sx.p, G.sy.p, G.pt.p, G.active.p, ns, tune->epsilon, G.part.p, G.stream,
                                       use_runs, lanes, use_class ? &G.cls : nullptr, false);
                if (rs) return rs;
                HIPCHK(hipEventRecord(G.ev[p * 2 * BATCH + 2 * i + 1], G.stream));
                HIPCHK(klaunch::step(G.st.p, G.sx.p, G.sy.p, G.pt.p, G.active.p, ns, G.part.p, nseg_fit, dpairs.p,
                                     n_pairs, dorder_p, dnext.p, dout.p, dt, dlive.p, PairQueue{},
                                     G.hist.p + p * BATCH + i, G.stream, G.rc));
            }
        launched += BATCH;
        for (int g = 0; g < ngroups; g++) {
            Group& G = grp[g];
            if (!G.queued[p]) continue;
            HIPCHK(hipMemcpyAsync(G.hist_h.data() + p * BATCH, G.hist.p + p * BATCH, BATCH * sizeof(int),
                                  hipMemcpyDeviceToHost, G.stream));
            HIPCHK(hipEventRecord(G.fin[p], G.stream));
        }
        return KESTREL_OK;
    };
    // waits for the batch in buffer p and counts (and times) its sweeps that
    // ran with live slots
    auto process = [&](int p) -> int {
        for (int g = 0; g < ngroups; g++)
            if (grp[g].queued[p]) HIPCHK(hipEventSynchronize(grp[g].fin[p]));
        for (int g = 0; g < ngroups; g++) {
            Group& G = grp[g];
            if (!G.queued[p]) continue;
            const int* hist = G.hist_h.data() + p * BATCH;
            const hipEvent_t* ev = G.ev.data() + p * 2 * BATCH;
            for (int i = 0; i < BATCH; i++) {
                if (G.live_before > 0 && (!use_graph || i == 0)) {
                    // with a graph, one interval covers the whole batch
                    // (sweeps, folds and steps)
                    float ms = 0, t_start = 0, t_end = 0;
                    const int j = use_graph ? 0 : 2 * i;
                    HIPCHK(hipEventElapsedTime(&ms, ev[j], ev[j + 1]));
                    HIPCHK(hipEventElapsedTime(&t_start, ref, ev[j]));
                    HIPCHK(hipEventElapsedTime(&t_end, ref, ev[j + 1]));
                    sweep_ms += ms;
                    spans.emplace_back(t_start, t_end);
                    end_ms = std::max<double>(end_ms, t_end);
                }
                if (G.live_before > 0) {
                    sweeps++;
                    // a slot goes idle only when no pair is left to claim
                    if (hist[i] < (int)ns) {
                        float t_end = 0;
                        HIPCHK(hipEventElapsedTime(&t_end, ref, ev[use_graph ? 1 : 2 * i + 1]));
                        if (t_end < exhaust_ms) exhaust_ms = t_end;
                    }
                }
                G.live_before = hist[i];
            }
        }
        return KESTREL_OK;
    };
    {
        const int rs = enqueue(0);
        if (rs) return rs;
    }
    for (int k = 0;; k++) {
        if (pipe && any_live()) {
            const int rs = enqueue((k + 1) & 1);
            if (rs) return rs;
        }
        int rs = process(pipe ? k & 1 : 0);
        if (rs) return rs;
        if (!any_live()) break;
        if (!pipe && (rs = enqueue(0))) return rs;
    }
    // a batch queued beyond the last live one
    for (int g = 0; g < ngroups; g++) HIPCHK(hipStreamSynchronize(grp[g].stream));
    // device time with at least one sweep running (the launches of the two
    // groups overlap)
    std::sort(spans.begin(), spans.end());
    double busy = 0, cur_s = -1, cur_e = -1;
    for (const auto& sp : spans) {
        if (sp.first > cur_e) {
            if (cur_e > cur_s) busy += cur_e - cur_s;
            cur_s = sp.first;
            cur_e = sp.second;
        } else if (sp.second > cur_e) {
            cur_e = sp.second;
        }
    }
    if (cur_e > cur_s) busy += cur_e - cur_s;
    c->last_busy_ms = busy;
    c->last_groups = ngroups;
    c->last_form = lanes ? KESTREL_FORM_LANES : use_class ? KESTREL_FORM_CLASS : KESTREL_FORM_TABLE;
    c->last_slots = ns;
    c->last_exhaust_ms = std::isfinite(exhaust_ms) ? exhaust_ms : end_ms;
    c->last_end_ms = end_ms;
    HIPCHK(hipMemcpyAsync(out, dout.p, n_pairs * sizeof(kestrel_output), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->last_recond_ms = 0;
    c->last_recond_pairs = 0;
    if (c->mixed) {
        const auto r0 = std::chrono::steady_clock::now();
        const int rs = recondition_records(
            c, c->geno.p, c->row_words, c->seg.p, c->nseg, *tune, out, n_pairs,
            [&](size_t i, uint32_t& x, uint32_t& y) {
                x = out[i].x;
                y = out[i].y;
            },
            st, &c->last_recond_pairs);
        if (rs) return rs;
        c->last_recond_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
    }
    c->last_sweep_ms = sweep_ms;
    c->last_sweeps = sweeps;
    c->last_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return KESTREL_OK;
}

int kestrel_autotune(kestrel_cohort* c, const kestrel_tune* tune, const uint32_t* pairs, size_t n_pairs,
                     uint64_t job_pairs, kestrel_autotune_result* res) {
    if (!c || !tune || !pairs || !res) return fail(KESTREL_EINVAL, "null argument");
    if (c->gt) return fail(KESTREL_EINVAL, "a GT cohort is fitted with kestrel_fit_pairs_gt");
    if (c->serve) return fail(KESTREL_EINVAL, "stop the serving engine before tuning");
    if (n_pairs < SWEEP_TPB) return fail(KESTREL_EINVAL, "need at least %d sample pairs", SWEEP_TPB);
    if (job_pairs == 0) job_pairs = n_pairs;
    const int64_t keep_slots = c->opt_slots, keep_streams = c->opt_streams, keep_batch = c->opt_batch;
    // the shapes tuned are those of the one-pair-per-lane sweep, whatever the
    // sample's size
    const int64_t keep_lanes = c->opt_pair_lanes;
    c->opt_pair_lanes = 1;
    std::vector<kestrel_output> out(n_pairs);
    uint32_t trials = 0;
    // Predicted milliseconds of a job of job_pairs pairs with one shape, from
    // the better of two fits of the sample: while pairs are left to claim
    // every slot is busy, so the sample's pairs completed by the time the
    // first slot goes idle over that time is the shape's steady rate; the rest
    // of the sample's fit (the drain of the last pairs in flight, and the
    // fixed host work around the device time) is added as measured.
    auto trial = [&](int64_t slots, int64_t streams, int64_t batch, double* ms) {
        c->opt_slots = slots;
        c->opt_streams = streams;
        c->opt_batch = batch;
        *ms = INFINITY;
        for (int r = 0; r < 2; r++) {
            const int rc = kestrel_fit_pairs(c, tune, pairs, n_pairs, 0, out.data());
            if (rc) return rc;
            trials++;
            const double inflight = (double)c->last_slots * c->last_groups;
            const double done = (double)n_pairs - inflight;
            double predicted = c->last_wall_ms;
            if (done > 0 && c->last_exhaust_ms > 0) {
                const double rate = done / c->last_exhaust_ms;  // pairs per ms while every slot is busy
                predicted = std::max(0.0, (double)job_pairs - inflight) / rate +
                            (c->last_wall_ms - c->last_exhaust_ms);
            }
            *ms = std::min(*ms, predicted);
        }
        return KESTREL_OK;
    };
    auto restore = [&] {
        c->opt_pair_lanes = keep_lanes;
        c->opt_slots = keep_slots;
        c->opt_streams = keep_streams;
        c->opt_batch = keep_batch;
    };
    double ms0 = 0;
    int rc = trial(keep_slots, keep_streams, keep_batch, &ms0);  // also warms the device up
    if (!rc) rc = trial(keep_slots, keep_streams, keep_batch, &ms0);
    if (rc) {
        restore();
        return rc;
    }
    // candidates: 1/8 to 3 device fills per group, one or two groups, 4 / 8 /
    // 16 iterations per synchronisation; a shape is tried only when the
    // sample holds at least 4 times its pairs in flight
    const uint64_t base = c->class_form() ? class_slots(c) / 2 : default_slots(c);
    double best_ms = ms0;
    int64_t best[3] = {keep_slots, keep_streams, keep_batch};
    for (const double fill : {0.125, 0.25, 0.5, 1.0, 2.0, 3.0}) {
        const int64_t slots = (int64_t)std::max<uint64_t>(
            SWEEP_TPB, ((uint64_t)(fill * (double)base) + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB);
        for (const int64_t streams : {1, 2}) {
            if ((uint64_t)slots * (uint64_t)streams * 4 > n_pairs) continue;
            for (const int64_t batch : {4, 8, 16}) {
                double ms = 0;
                rc = trial(slots, streams, batch, &ms);
                if (rc) {
                    restore();
                    return rc;
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best[0] = slots;
                    best[1] = streams;
                    best[2] = batch;
                }
            }
        }
    }
    c->opt_pair_lanes = keep_lanes;
    c->opt_slots = best[0];
    c->opt_streams = best[1];
    c->opt_batch = best[2];
    res->slots = best[0];
    res->streams = best[1];
    res->batch = best[2];
    res->pairs_per_s = (double)job_pairs / (best_ms * 1e-3);
    res->default_pairs_per_s = (double)job_pairs / (ms0 * 1e-3);
    res->trials = trials;
    return KESTREL_OK;
}

int kestrel_evaluate(kestrel_cohort* c, const uint32_t* pairs, const double* points, size_t n, double eps,
                     double* f, double* g, double* h, uint32_t* n_sites) {
    if (!c || !pairs || !points || !f || !g || !h || !n_sites) return fail(KESTREL_EINVAL, "null argument");
    if (c->gt_parts[0] || c->gt_parts[1])
        return fail(KESTREL_EINVAL, "a -f GT cohort with 3- or 4-allele sites has no likelihood evaluation");
    if (n == 0) return KESTREL_OK;
    for (size_t i = 0; i < 2 * n; i++)
        if (pairs[i] >= c->n_samples) return fail(KESTREL_EINVAL, "pair names sample %u of %u", pairs[i], c->n_samples);
    const uint32_t ns = (uint32_t)((n + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB);
    std::vector<uint32_t> hx(ns, 0), hy(ns, 0);
    std::vector<int> hact(ns, 0), hex(ns, 0);
    std::vector<double> hpt((size_t)9 * ns, 0.0);
    bool any_exact = false;
    for (size_t i = 0; i < n; i++) {
        hx[i] = pairs[2 * i];
        hy[i] = pairs[2 * i + 1];
        // points outside the sweeps' domain go to the exact kernel
        const bool in = point_in_domain(points + i * 9, eps, c->lmax);
        hact[i] = in ? 1 : 0;
        hex[i] = in ? 0 : 1;
        any_exact |= !in;
        for (int k = 0; k < 9; k++) hpt[(size_t)k * ns + i] = points[i * 9 + k];
    }
    DevBuf<uint32_t> dsx, dsy, dn;
    DevBuf<double> dpt, dpart, df, dg, dh;
    DevBuf<int> dact, dex;
    HIPCHK(dsx.alloc(ns));
    HIPCHK(dsy.alloc(ns));
    HIPCHK(dact.alloc(ns));
    HIPCHK(dex.alloc(ns));
    HIPCHK(dpt.alloc((size_t)9 * ns));
    // the class-sorted sweep writes its one segment, the exact kernel the
    // cohort's segments: partials start at zero and are folded over both
    const uint64_t class_cap = c->class_form() ? class_slot_cap(c->class_blocks(), 1, c->opt_class_list_mib) : 0;
    if (c->class_form() && class_cap == 0)
        return fail(KESTREL_EDEVICE, "device memory for the class-sorted sweep's lists of 256 points (%llu bytes) is "
                    "not free (option class_sweep 0 selects the site-table sweep)",
                    (unsigned long long)(256 * ClassLists::slot_bytes(c->class_blocks())));
    const bool use_class = c->class_form();
    const uint32_t nfold = std::max(c->nseg_total(), use_class ? c->nseg_class_total() : 0u);
    HIPCHK(dpart.alloc((size_t)nfold * NPART * ns));
    HIPCHK(df.alloc(n));
    HIPCHK(dg.alloc(9 * n));
    HIPCHK(dh.alloc(81 * n));
    HIPCHK(dn.alloc(n));
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(dsx.p, hx.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dsy.p, hy.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dact.p, hact.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dpt.p, hpt.data(), hpt.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dex.p, hex.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
    ClassLists cls;
    if (use_class) {
        HIPCHK(hipMemsetAsync(dpart.p, 0, dpart.n * sizeof(double), st));
        // the lists of at most class_cap points at a time
        const int rc = cls.alloc((uint32_t)std::min<uint64_t>(ns, class_cap), c->class_blocks(), st);
        if (rc) return rc;
    }
    int rs = launch_sweeps(c, dsx.p, dsy.p, dpt.p, dact.p, ns, eps, dpart.p, st, false, false,
                           use_class ? &cls : nullptr);
    if (rs) return rs;
    if (any_exact) {
        rs = launch_exact(c, dsx.p, dsy.p, dpt.p, dex.p, ns, eps, dpart.p, st);
        if (rs) return rs;
    }
    HIPCHK(klaunch::finalize(dpart.p, nfold, ns, (uint32_t)n, df.p, dg.p, dh.p, dn.p, st));
    HIPCHK(hipMemcpyAsync(f, df.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(g, dg.p, 9 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(h, dh.p, 81 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(n_sites, dn.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return KESTREL_OK;
}

void kestrel_cohort_destroy(kestrel_cohort* c) { delete c; }

}  // extern "C"

namespace {

// Times one launch of the FMA-chain kernel (after a short warm-up launch).
template <class T, class Launch>
int measure_fma(Launch launch, double* tflops) {
    if (!tflops) return fail(KESTREL_EINVAL, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KESTREL_EDEVICE, "no HIP device: libkestrel_amd needs an MI355X (gfx950) GPU");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 16;  // 16 workgroups of 4 waves per CU
    const uint32_t iters = 1u << 15;
    DevBuf<T> sink;
    HIPCHK(sink.alloc(1));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(launch(blocks, 1024u, sink.p));  // warm up clocks
    HIPCHK(hipEventRecord(e0, nullptr));
    HIPCHK(launch(blocks, iters, sink.p));
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double flop = (double)blocks * 256.0 * iters * 16.0 * 2.0;
    *tflops = flop / (ms * 1e-3) / 1e12;
    return KESTREL_OK;
}

}  // namespace

extern "C" {

int kestrel_measure_fp64_peak(double* tflops) {
    return measure_fma<double>(
        [](uint32_t b, uint32_t it, double* s) { return klaunch::calib_fp64(b, it, s, nullptr); }, tflops);
}

int kestrel_measure_fp32_peak(double* tflops) {
    return measure_fma<float>(
        [](uint32_t b, uint32_t it, float* s) { return klaunch::calib_fp32(b, it, s, nullptr); }, tflops);
}

int kestrel_measure_lds(double* tbs) {
    if (!tbs) return fail(KESTREL_EINVAL, "null argument");
    // the FMA timer's launch shape; 8 loads of 16 bytes per lane and step
    double per = 0;
    const int rc = measure_fma<double>(
        [](uint32_t b, uint32_t it, double* s) { return klaunch::calib_lds(b, it, s, nullptr); }, &per);
    if (rc) return rc;
    // measure_fma counts 16 chains x 2 FLOP per lane and step; this kernel moves 8 x 16 bytes
    *tbs = per * (8.0 * 16.0) / (16.0 * 2.0);
    return KESTREL_OK;
}

int kestrel_measure_copy(uint64_t bytes, double* gbps) {
    if (!gbps || bytes < 16) return fail(KESTREL_EINVAL, "bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KESTREL_EDEVICE, "no HIP device: libkestrel_amd needs an MI355X (gfx950) GPU");
    const uint64_t n16 = bytes / 16;
    DevBuf<uint4> a, b;
    HIPCHK(a.alloc(n16));
    HIPCHK(b.alloc(n16));
    HIPCHK(hipMemset(a.p, 1, n16 * 16));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(klaunch::copy(a.p, b.p, n16, nullptr));
    HIPCHK(hipEventRecord(e0, nullptr));
    HIPCHK(klaunch::copy(a.p, b.p, n16, nullptr));
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbps = 2.0 * (double)(n16 * 16) / (ms * 1e-3) / 1e9;
    return KESTREL_OK;
}

int kestrel_selftest(uint64_t n, uint64_t seed, uint64_t* mismatches) {
    if (!mismatches) return fail(KESTREL_EINVAL, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KESTREL_EDEVICE, "no HIP device: libkestrel_amd needs an MI355X (gfx950) GPU");
    DevBuf<unsigned long long> bad;
    HIPCHK(bad.alloc(1));
    HIPCHK(hipMemset(bad.p, 0, sizeof(unsigned long long)));
    HIPCHK(klaunch::selftest(n, seed, bad.p, nullptr));
    unsigned long long h = 0;
    HIPCHK(hipMemcpy(&h, bad.p, sizeof h, hipMemcpyDeviceToHost));
    *mismatches = h;
    return KESTREL_OK;
}

}  // extern "C"

namespace {

// The 2-allele sites of a -f GT cohort (wrapper.rs:22-87): sites with a
// missing call in any sample are dropped (vcf.rs:217-221), frequencies are
// counted (wrapper.rs:412-428) unless given (n_sites x 2).
int init_gt2(kestrel_cohort* c, const uint8_t* packed, uint32_t n_sites, size_t row_stride_bytes, const double* af) {
    const uint32_t n_samples = c->n_samples;
    c->n_sites = n_sites;
    DevBuf<uint8_t> dpacked, dflag;
    HIPCHK(dpacked.alloc((size_t)n_samples * row_stride_bytes));
    HIPCHK(dflag.alloc(n_sites));
    if (n_sites)
        HIPCHK(hipMemcpyAsync(dpacked.p, packed, (size_t)n_samples * row_stride_bytes, hipMemcpyHostToDevice,
                              c->stream));
    HIPCHK(klaunch::complete_sites(dpacked.p, row_stride_bytes, n_samples, n_sites, dflag.p, c->stream));
    c->kept_mask.resize(n_sites);
    if (n_sites)
        HIPCHK(hipMemcpyAsync(c->kept_mask.data(), dflag.p, n_sites, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> kept;
    for (uint32_t s = 0; s < n_sites; s++)
        if (c->kept_mask[s]) kept.push_back(s);
    int rc = repack_kept(c, dpacked.p, row_stride_bytes, kept);
    if (rc) return rc;
    HIPCHK(c->af.alloc((size_t)c->v_pad * 2));
    HIPCHK(hipMemsetAsync(c->af.p, 0, (size_t)c->v_pad * 2 * sizeof(double), c->stream));
    c->af_all.assign((size_t)n_sites * 2, std::nan(""));
    if (af) {
        std::vector<double> afk((size_t)c->v_pad * 2, 0.0);
        for (uint32_t t = 0; t < c->n_kept; t++) {
            afk[2 * t] = af[2 * kept[t]];
            afk[2 * t + 1] = af[2 * kept[t] + 1];
        }
        HIPCHK(hipMemcpyAsync(c->af.p, afk.data(), afk.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(klaunch::gt_count(c->geno.p, c->row_words, n_samples, c->n_kept, c->af.p, c->stream));
    }
    std::vector<double> afk((size_t)c->v_pad * 2);
    HIPCHK(hipMemcpyAsync(afk.data(), c->af.p, afk.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t t = 0; t < c->n_kept; t++) {
        c->af_all[2 * kept[t]] = afk[2 * t];
        c->af_all[2 * kept[t] + 1] = afk[2 * t + 1];
    }
    return c->set_segments();
}

// The 3- or 4-allele sites of a -f GT cohort: idx n_samples x n_in genotype
// indices (VCF order, 255 = missing call); sites with a missing call are
// dropped, the rest packed into 4-bit rows; frequencies counted on the device
// unless given (n_in x A); then the sites' m tables.  af_all receives the
// frequencies of every input site (NaN for a dropped one).
int init_gt_part(kestrel_cohort* c, int A, const uint8_t* idx, uint32_t n_in, const double* af,
                 std::vector<double>& af_all, uint32_t* n_kept) {
    const uint32_t n = c->n_samples, G = (uint32_t)(A * (A + 1) / 2);
    std::vector<uint8_t> complete(n_in, 1);
    for (uint32_t i = 0; i < n; i++) {  // row by row (sample-major input)
        const uint8_t* row = idx + (size_t)i * n_in;
        for (uint32_t s = 0; s < n_in; s++) {
            const uint8_t g = row[s];
            if (g == 255) complete[s] = 0;
            else if (g >= G)
                return fail(KESTREL_EINVAL, "genotype index %u of sample %u at %u-allele site %u exceeds %u", g, i, A,
                            s, G - 1);
        }
    }
    std::vector<uint32_t> kept;
    for (uint32_t s = 0; s < n_in; s++)
        if (complete[s]) kept.push_back(s);
    *n_kept = (uint32_t)kept.size();
    af_all.assign((size_t)n_in * A, std::nan(""));
    if (kept.empty()) return KESTREL_OK;
    auto part = std::make_unique<GtPart>();
    part->alleles = A;
    part->n_sites = (uint32_t)kept.size();
    part->row_words = (part->n_sites + 7) / 8;
    std::vector<uint32_t> rows((size_t)n * part->row_words, 0xffffffffu);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t* row = rows.data() + (size_t)i * part->row_words;
        for (uint32_t t = 0; t < part->n_sites; t++) {
            const uint32_t g = idx[(size_t)i * n_in + kept[t]], sh = 4 * (t % 8);
            row[t / 8] = (row[t / 8] & ~(15u << sh)) | (g << sh);
        }
    }
    HIPCHK(part->geno.alloc(rows.size()));
    HIPCHK(hipMemcpyAsync(part->geno.p, rows.data(), rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(part->af.alloc((size_t)part->n_sites * A));
    std::vector<double> afk((size_t)part->n_sites * A);
    if (af) {
        for (uint32_t t = 0; t < part->n_sites; t++)
            for (int a = 0; a < A; a++) afk[(size_t)t * A + a] = af[(size_t)kept[t] * A + a];
        HIPCHK(hipMemcpyAsync(part->af.p, afk.data(), afk.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(klaunch::gt_count_multi(A, part->geno.p, part->row_words, n, part->n_sites, part->af.p, c->stream));
    }
    HIPCHK(part->mtab.alloc((size_t)part->n_sites * G * G * 9));
    HIPCHK(klaunch::cls_m(A, part->af.p, part->n_sites, part->mtab.p, c->stream));
    HIPCHK(hipMemcpyAsync(afk.data(), part->af.p, afk.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t t = 0; t < part->n_sites; t++)
        for (int a = 0; a < A; a++) af_all[(size_t)kept[t] * A + a] = afk[(size_t)t * A + a];
    c->gt_parts[A - 3] = std::move(part);
    return KESTREL_OK;
}

bool all_finite(const double* v, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

extern "C" {

int kestrel_cohort_create_gt(kestrel_cohort** out, const uint8_t* packed, uint32_t n_samples, uint32_t n_sites,
                             size_t row_stride_bytes, const double* af) {
    if (!out || !packed) return fail(KESTREL_EINVAL, "null argument");
    *out = nullptr;
    if (n_samples == 0 || n_sites == 0) return fail(KESTREL_EINVAL, "need at least one sample and one site");
    if (row_stride_bytes < (n_sites + 3) / 4) return fail(KESTREL_EINVAL, "row stride too small");
    if (af && !all_finite(af, (size_t)n_sites * 2)) return fail(KESTREL_EINVAL, "allele frequencies must be finite");
    auto c = new kestrel_cohort();
    std::unique_ptr<kestrel_cohort> guard(c);
    int rc = open_device(c);
    if (rc) return rc;
    c->gt = true;
    c->n_samples = n_samples;
    rc = init_gt2(c, packed, n_sites, row_stride_bytes, af);
    if (rc) return rc;
    *out = guard.release();
    return KESTREL_OK;
}

int kestrel_cohort_create_gt_multi(kestrel_cohort** out, uint32_t n_samples, const uint8_t* packed2, uint32_t n2,
                                   size_t row_stride_bytes, const uint8_t* gt3, uint32_t n3, const uint8_t* gt4,
                                   uint32_t n4, const double* af2, const double* af3, const double* af4) {
    if (!out) return fail(KESTREL_EINVAL, "null argument");
    *out = nullptr;
    if (n_samples == 0 || (uint64_t)n2 + n3 + n4 == 0) return fail(KESTREL_EINVAL, "need at least one sample and one site");
    if ((n2 && !packed2) || (n3 && !gt3) || (n4 && !gt4)) return fail(KESTREL_EINVAL, "null genotype bucket");
    if (n2 && row_stride_bytes < (n2 + 3) / 4) return fail(KESTREL_EINVAL, "row stride too small");
    const bool given = af2 || af3 || af4;
    if (given && ((n2 && !af2) || (n3 && !af3) || (n4 && !af4)))
        return fail(KESTREL_EINVAL, "allele frequencies must be given for every bucket or none");
    if (given && ((af2 && !all_finite(af2, (size_t)n2 * 2)) || (af3 && !all_finite(af3, (size_t)n3 * 3)) ||
                  (af4 && !all_finite(af4, (size_t)n4 * 4))))
        return fail(KESTREL_EINVAL, "allele frequencies must be finite");
    auto c = new kestrel_cohort();
    std::unique_ptr<kestrel_cohort> guard(c);
    int rc = open_device(c);
    if (rc) return rc;
    c->gt = true;
    c->multi = true;
    c->n_samples = n_samples;
    rc = init_gt2(c, packed2, n2, n2 ? row_stride_bytes : 0, af2);
    if (rc) return rc;
    c->part_sites[0] = c->n_kept;
    c->af_in[0] = c->af_all;
    rc = init_gt_part(c, 3, gt3, n3, af3, c->af_in[1], &c->part_sites[1]);
    if (rc) return rc;
    rc = init_gt_part(c, 4, gt4, n4, af4, c->af_in[2], &c->part_sites[2]);
    if (rc) return rc;
    *out = guard.release();
    return KESTREL_OK;
}

int kestrel_fit_pairs_gt(kestrel_cohort* c, uint64_t iterations, const uint32_t* pairs, size_t n_pairs,
                         uint32_t n_select, kestrel_output* out, double* q_out) {
    if (!c || !out) return fail(KESTREL_EINVAL, "null argument");
    if (!c->gt) return fail(KESTREL_EINVAL, "cohort was not created by kestrel_cohort_create_gt");
    if (iterations == 0) return fail(KESTREL_EINVAL, "iterations must be non-zero");
    const GtPart* p3 = c->gt_parts[0].get();
    const GtPart* p4 = c->gt_parts[1].get();
    const uint32_t n3 = p3 ? p3->n_sites : 0, n4 = p4 ? p4->n_sites : 0;
    const uint32_t v_total = c->n_kept + n3 + n4;
    if (v_total == 0) return fail(KESTREL_EINVAL, "Must have at least one site");
    std::vector<uint32_t> tri;
    if (!pairs) {
        if (n_select > c->n_samples)
            return fail(KESTREL_EINVAL, "Requested samples %u exceeds total samples %u", n_select, c->n_samples);
        n_pairs = (size_t)n_select * ((size_t)n_select + 1) / 2;
        tri.resize(2 * n_pairs);
        size_t i = 0;
        for (uint32_t x = 0; x < n_select; x++)
            for (uint32_t y = x; y < n_select; y++, i++) {
                tri[2 * i] = x;
                tri[2 * i + 1] = y;
            }
        pairs = tri.data();
    } else {
        for (size_t i = 0; i < 2 * n_pairs; i++)
            if (pairs[i] >= c->n_samples) return fail(KESTREL_EINVAL, "pair names sample %u of %u", pairs[i], c->n_samples);
    }
    hipStream_t st = c->stream;
    if (!c->have_q) {
        DevBuf<double> z;
        HIPCHK(z.alloc(std::max<uint64_t>(1, klaunch::gram_tiles(c->n_kept, n3, n4)) * 45 * 8));
        HIPCHK(c->qmat.alloc(81));
        HIPCHK(klaunch::gram(c->af.p, c->n_kept, p3 ? p3->mtab.p : nullptr, n3, p4 ? p4->mtab.p : nullptr, n4, z.p,
                             c->qmat.p, st));
        HIPCHK(hipStreamSynchronize(st));
        c->have_q = true;
    }
    if (q_out) HIPCHK(hipMemcpy(q_out, c->qmat.p, 81 * sizeof(double), hipMemcpyDeviceToHost));
    if (n_pairs == 0) return KESTREL_OK;
    const auto t0 = std::chrono::steady_clock::now();
    kestrel_tune t;
    kestrel_tune_default(&t, iterations);
    t.qp_max_iter = iterations;  // coefficients.rs:207: Tuneables::new(iterations, iterations)
    DevBuf<uint32_t> dpairs;
    DevBuf<kestrel_output> dout;
    DevBuf<double> c34;
    HIPCHK(dpairs.alloc(2 * n_pairs));
    HIPCHK(dout.alloc(n_pairs));
    HIPCHK(hipMemcpyAsync(dpairs.p, pairs, 2 * n_pairs * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    struct EvGuard {
        hipEvent_t& a;
        hipEvent_t& b;
        ~EvGuard() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evg{e0, e1};
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    if (p3 || p4) {
        HIPCHK(c34.alloc(n_pairs * 9));
        if (p3)
            HIPCHK(klaunch::cls_multi(3, p3->geno.p, p3->row_words, p3->mtab.p, n3, dpairs.p, n_pairs, 0, c34.p, st));
        if (p4)
            HIPCHK(klaunch::cls_multi(4, p4->geno.p, p4->row_words, p4->mtab.p, n4, dpairs.p, n_pairs, p3 ? 1 : 0,
                                      c34.p, st));
    }
    DevBuf<double> csum;
    HIPCHK(csum.alloc(n_pairs * 9));
    HIPCHK(klaunch::cls_pairs(c->geno.p, c->row_words, c->af.p, c->n_kept, c34.p, v_total, c->qmat.p, dpairs.p,
                              n_pairs, to_dev(t), csum.p, dout.p, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipMemcpyAsync(out, dout.p, n_pairs * sizeof(kestrel_output), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float kms = 0;
    HIPCHK(hipEventElapsedTime(&kms, e0, e1));
    c->last_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->last_sweeps = 1;
    c->last_sweep_ms = kms;  // the per-pair CLS kernels (c3/c4 and cls_pairs_kernel)
    c->last_busy_ms = kms;
    c->last_groups = 1;
    return KESTREL_OK;
}

}  // extern "C"
