// serve.cpp -- the serving engine: a device-resident cohort answering
// requests of query samples with continuous batching.
//
// Device state of an engine:
//   geno      the cohort's kept-site rows followed by max_query_rows rows for
//             queries (rows are lent to admitted requests)
//   ring      pair entries (x row, y row, output record, request slot); the
//             host appends entries of the admitted requests (shortest
//             remaining first, or round robin) a bounded backlog at a time,
//             so a small request is not queued behind a large one
//   outpool   max_pairs output records, a contiguous range per request
//   ctrl      published tail, claim counter, busy slots, done count per
//             request slot
// The engine thread repeats: admit, append entries, sweep, step, read ctrl
// back; a request whose done count reaches its pair count is copied out.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "host.hpp"

namespace {

using Clock = std::chrono::steady_clock;

enum : int { RQ_PENDING = 0, RQ_LIVE = 1, RQ_DONE = 2, RQ_FAILED = 3 };

constexpr uint32_t MAX_REQ = 1024;     // request slots (concurrently admitted)
constexpr uint64_t FEED_CHUNK = 256;   // entries per request per round-robin turn
constexpr size_t CTRL_DONE_OFF = 32;   // byte offset of done[] in ctrl
constexpr double SHORT_S = 1.0;        // a short request: the device finishes it within this many seconds alone

struct Request {
    uint64_t id = 0;
    uint32_t n_query = 0, n_sub = 0;
    std::vector<uint8_t> qdata;     // n_query rows of rb bytes (dropped at admission)
    std::vector<uint32_t> subset;   // empty: the whole cohort
    uint64_t npairs = 0;
    kestrel_output* out = nullptr;
    int state = RQ_PENDING;
    std::string err;
    uint32_t slot = 0;
    std::vector<uint32_t> rows;     // geno rows of the queries
    uint64_t out_off = 0;
    uint64_t fed = 0;               // entries appended to the ring
    Clock::time_point t_submit, t_done;
};

// First-fit allocator of output-record ranges.
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    void init(uint64_t n) {
        free_.clear();
        if (n) free_[0] = n;
    }
    bool take(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, o = it->first + n;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        return false;
    }
    void give(uint64_t off, uint64_t n) {
        if (!n) return;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace

struct ServeState {
    kestrel_cohort* c = nullptr;
    kestrel_tune tune{};
    kdev::Tune dt{};
    hipStream_t stream = nullptr;
    uint32_t qmax = 0, n_cohort = 0, row_words = 0, nseg = 0, ns = 0;
    bool cls = false;  // sweeps by the class-sorted form (partials in one segment)
    uint64_t out_cap = 0, ring_cap = 0, backlog = 0;
    int policy = 1;  // 0 round robin, 1 shortest remaining first (with aging)
    double rate = 0;                 // pairs per second the device consumes (moving average)
    Clock::time_point t_round{};     // start of the previous engine round
    size_t rb = 0;  // bytes of one query row over n_sites

    DevBuf<uint32_t> geno, kept, seg, qrows, ring;
    DevBuf<uint8_t> qstage, ctrl;
    DevBuf<kestrel_output> outpool;
    // slot groups: each sweeps and steps its own slots on its own stream,
    // all claim from the one ring (as the slot groups of a batch fit)
    struct Group {
        hipStream_t stream = nullptr;  // group 0 uses ServeState::stream
        DevBuf<uint32_t> sx, sy;
        DevBuf<SlotState> slots;
        DevBuf<double> pt, partial;
        DevBuf<int> active;
        ClassLists cls;  // the class-sorted sweep's pattern lists (FP64 cohorts)
    };
    Group grp[2];
    int ngroups = 1;
    int iters = 2;                 // sweep + step iterations per engine round
    hipStream_t stream2 = nullptr;
    hipEvent_t fed = nullptr, swept = nullptr;
    uint8_t* hctrl = nullptr;     // pinned read-back of ctrl
    uint32_t* hent = nullptr;     // pinned staging of appended entries
    uint64_t device_bytes = 0;

    std::mutex mu;
    std::condition_variable cv_engine, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> live;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<uint32_t> free_slots, free_rows;
    RangeAlloc outs;
    uint64_t next_id = 1;
    bool stop = false, failed = false;
    std::string fail_msg;
    std::thread th;

    // engine-thread state
    uint64_t tail = 0, consumed = 0;
    int busy = 0;
    size_t rr = 0;  // round-robin position in live
    // statistics (under mu)
    uint64_t st_requests = 0, st_pairs = 0, st_sweeps = 0, st_busy = 0;

    ~ServeState() {
        if (hctrl) (void)hipHostFree(hctrl);
        if (hent) (void)hipHostFree(hent);
        if (fed) (void)hipEventDestroy(fed);
        if (swept) (void)hipEventDestroy(swept);
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
    }
};

namespace {

unsigned long long* ctrl_tail(ServeState* S) { return reinterpret_cast<unsigned long long*>(S->ctrl.p); }
unsigned long long* ctrl_next(ServeState* S) { return reinterpret_cast<unsigned long long*>(S->ctrl.p + 8); }
int* ctrl_busy(ServeState* S) { return reinterpret_cast<int*>(S->ctrl.p + 16); }
uint32_t* ctrl_done(ServeState* S) { return reinterpret_cast<uint32_t*>(S->ctrl.p + CTRL_DONE_OFF); }

// Resources of a pending request (under mu); false leaves nothing taken.
bool try_admit(ServeState* S, Request* r) {
    if (S->free_slots.empty() || S->free_rows.size() < r->n_query) return false;
    uint64_t off = 0;
    if (!S->outs.take(r->npairs, &off)) return false;
    r->out_off = off;
    r->slot = S->free_slots.back();
    S->free_slots.pop_back();
    r->rows.assign(S->free_rows.end() - r->n_query, S->free_rows.end());
    S->free_rows.resize(S->free_rows.size() - r->n_query);
    r->state = RQ_LIVE;
    return true;
}

void release(ServeState* S, Request* r) {
    S->outs.give(r->out_off, r->npairs);
    S->free_slots.push_back(r->slot);
    S->free_rows.insert(S->free_rows.end(), r->rows.begin(), r->rows.end());
}

// Upload an admitted request's query rows into its geno rows and clear its
// done counter (engine thread; ordered on the engine stream).
int upload(ServeState* S, Request* r) {
    HIPCHK(hipMemcpyAsync(S->qstage.p, r->qdata.data(), r->qdata.size(), hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipMemcpyAsync(S->qrows.p, r->rows.data(), r->rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          S->stream));
    HIPCHK(klaunch::gather(S->qstage.p, S->rb, S->kept.p, S->c->n_kept, r->n_query, S->geno.p, S->row_words,
                           S->stream, S->qrows.p));
    HIPCHK(hipMemsetAsync(ctrl_done(S) + r->slot, 0, sizeof(uint32_t), S->stream));
    // query rows may have been rewritten: every slot's pattern lists are built
    // again (both streams are idle between engine rounds)
    if (S->cls)
        for (int g = 0; g < S->ngroups; g++)
            HIPCHK(hipMemsetAsync(S->grp[g].cls.built.p, 0xff, (size_t)S->ns * sizeof(uint2), S->stream));
    // the staging buffers are reused by the next upload only after this one
    // has run
    HIPCHK(hipStreamSynchronize(S->stream));
    std::vector<uint8_t>().swap(r->qdata);
    return KESTREL_OK;
}

// Append entries until the unread backlog reaches S->backlog, then publish
// the new tail.  Round robin gives every live request FEED_CHUNK entries in
// turn (processor sharing); shortest remaining first feeds the request with
// the fewest unfed entries first (lower mean latency under load) in two
// classes: requests the device finishes within SHORT_S seconds alone go
// first by remaining entries; the others follow by remaining entries minus
// rate x seconds since submission (rate x the time the request would finish
// if fed alone from its arrival), so a large request is not starved by a
// stream of smaller ones.  Measured in a fluid model of the C5 request mix at
// 0.7 load: p99 18 s against 23 s for plain shortest-remaining-first and
// 30 s for round robin, 1-query requests unchanged.
int feed(ServeState* S, const std::vector<Request*>& live) {
    uint64_t room = S->backlog > S->tail - S->consumed ? S->backlog - (S->tail - S->consumed) : 0;
    if (room == 0 || live.empty()) return KESTREL_OK;
    uint64_t n = 0;
    const uint64_t t0 = S->tail;
    auto append = [&](Request* r, uint64_t k) {
        for (uint64_t i = r->fed; i < r->fed + k; i++, n++) {
            const uint64_t q = i / r->n_sub, j = i % r->n_sub;
            uint32_t* e = S->hent + 4 * n;
            e[0] = r->rows[q];
            e[1] = r->subset.empty() ? (uint32_t)j : r->subset[j];
            e[2] = (uint32_t)(r->out_off + i);
            e[3] = r->slot;
        }
        r->fed += k;
        room -= k;
    };
    if (S->policy == 1) {
        std::vector<Request*> order(live);
        const auto now = Clock::now();
        std::vector<std::pair<double, size_t>> key(order.size());
        constexpr double LONG_CLASS = 1e18;  // keys of the second class start here
        for (size_t i = 0; i < order.size(); i++) {
            const double waited = std::chrono::duration<double>(now - order[i]->t_submit).count();
            const double rem = (double)(order[i]->npairs - order[i]->fed);
            key[i] = {(double)order[i]->npairs <= S->rate * SHORT_S ? rem : LONG_CLASS + rem - S->rate * waited, i};
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<double, size_t>& a, const std::pair<double, size_t>& b) {
                             return a.first < b.first;
                         });
        for (size_t i = 0; i < key.size(); i++) order[i] = live[key[i].second];
        for (Request* r : order) {
            if (room == 0) break;
            append(r, std::min<uint64_t>(room, r->npairs - r->fed));
        }
    } else {
        size_t idle = 0;
        while (room > 0 && idle < live.size()) {
            Request* r = live[S->rr % live.size()];
            S->rr++;
            const uint64_t k = std::min<uint64_t>({FEED_CHUNK, room, r->npairs - r->fed});
            if (k == 0) {
                idle++;
                continue;
            }
            idle = 0;
            append(r, k);
        }
    }
    if (n == 0) return KESTREL_OK;
    // ring positions t0 .. t0+n-1 (mod ring_cap), in at most two copies
    const uint64_t p0 = t0 % S->ring_cap, first = std::min<uint64_t>(n, S->ring_cap - p0);
    HIPCHK(hipMemcpyAsync(S->ring.p + 4 * p0, S->hent, first * 16, hipMemcpyHostToDevice, S->stream));
    if (first < n)
        HIPCHK(hipMemcpyAsync(S->ring.p, S->hent + 4 * first, (n - first) * 16, hipMemcpyHostToDevice, S->stream));
    S->tail = t0 + n;
    reinterpret_cast<unsigned long long*>(S->hctrl)[0] = S->tail;
    HIPCHK(hipMemcpyAsync(ctrl_tail(S), S->hctrl, sizeof(unsigned long long), hipMemcpyHostToDevice, S->stream));
    return KESTREL_OK;
}

// Copy a finished request's records to the caller: x = query index.
int finish(ServeState* S, Request* r) {
    if (r->npairs) {
        HIPCHK(hipMemcpyAsync(r->out, S->outpool.p + r->out_off, r->npairs * sizeof(kestrel_output),
                              hipMemcpyDeviceToHost, S->stream));
        HIPCHK(hipStreamSynchronize(S->stream));
    }
    if (S->c->mixed) {  // condition estimates from FP64 where the FP32-summed Hessian cannot resolve them
        const int rc = recondition_records(
            S->c, S->geno.p, S->row_words, S->seg.p, S->nseg, S->tune, r->out, r->npairs,
            [&](size_t i, uint32_t& x, uint32_t& y) {
                const uint64_t j = i % r->n_sub;
                x = r->rows[i / r->n_sub];
                y = r->subset.empty() ? (uint32_t)j : r->subset[j];
            },
            S->stream);
        if (rc) return rc;
    }
    for (uint64_t i = 0; i < r->npairs; i++) r->out[i].x = (uint32_t)(i / r->n_sub);
    return KESTREL_OK;
}

// One engine round; returns a KESTREL_E* code on a device error.
int engine_round(ServeState* S, bool* idle) {
    std::vector<Request*> admitted, live;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_engine.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->live.empty(); });
        if (S->stop) {
            *idle = true;
            return KESTREL_OK;
        }
        while (!S->pending.empty() && try_admit(S, S->pending.front())) {
            admitted.push_back(S->pending.front());
            S->live.push_back(S->pending.front());
            S->pending.pop_front();
        }
        // behind a request that does not fit yet, short requests (those the
        // device finishes within a second alone) that do fit are admitted
        for (auto it = S->pending.begin(); it != S->pending.end();) {
            Request* r = *it;
            if ((double)r->npairs <= S->rate * SHORT_S && try_admit(S, r)) {
                admitted.push_back(r);
                S->live.push_back(r);
                it = S->pending.erase(it);
            } else {
                ++it;
            }
        }
        live = S->live;
    }
    *idle = false;
    for (Request* r : admitted) {
        int rc = upload(S, r);
        if (rc) return rc;
    }
    int rc = feed(S, live);
    if (rc) return rc;
    // the other group's stream starts after the ring and its tail are written
    if (S->ngroups == 2) {
        HIPCHK(hipEventRecord(S->fed, S->stream));
        HIPCHK(hipStreamWaitEvent(S->grp[1].stream, S->fed, 0));
    }
    const PairQueue q{S->ring.p, S->ring_cap, ctrl_tail(S), ctrl_done(S), ctrl_busy(S)};
    const bool sweep = S->busy > 0 || S->tail > S->consumed;
    for (int it = 0; it < S->iters; it++)
        for (int g = 0; g < S->ngroups; g++) {
            ServeState::Group& G = S->grp[g];
            if (sweep && S->cls) {
                HIPCHK(klaunch::class_build(S->geno.p, S->row_words, G.sx.p, G.sy.p, G.active.p, S->ns, 0, S->ns,
                                            G.cls.nblk, G.cls.lists.p, G.cls.counts.p, G.cls.built.p, G.stream));
                HIPCHK(klaunch::sweep_class(S->c->af.p, S->c->v_pad, S->c->seg_cls.p, S->c->nseg_cls, G.cls.lists.p,
                                            G.cls.counts.p, G.cls.nblk, G.pt.p, G.active.p, S->ns, 0, S->ns,
                                            S->tune.epsilon, G.partial.p, G.stream, (int)S->c->opt_class_waves));
            } else if (sweep) {
                HIPCHK(klaunch::sweep(S->geno.p, S->row_words, S->c->af.p, S->seg.p, S->nseg, G.sx.p, G.sy.p, G.pt.p,
                                      G.active.p, S->ns, S->tune.epsilon, G.partial.p, G.stream, nullptr, 0, 3,
                                      false, S->c->mixed));
            }
            HIPCHK(klaunch::step(G.slots.p, G.sx.p, G.sy.p, G.pt.p, G.active.p, S->ns, G.partial.p,
                                 S->cls ? S->c->nseg_cls : S->nseg, nullptr,
                                 0, nullptr, ctrl_next(S), S->outpool.p, S->dt, nullptr, q, nullptr, G.stream));
        }
    if (S->ngroups == 2) {
        HIPCHK(hipEventRecord(S->swept, S->grp[1].stream));
        HIPCHK(hipStreamWaitEvent(S->stream, S->swept, 0));
    }
    uint32_t max_slot = 0;
    for (Request* r : live) max_slot = std::max(max_slot, r->slot + 1);
    HIPCHK(hipMemcpyAsync(S->hctrl + 8, S->ctrl.p + 8, CTRL_DONE_OFF - 8 + 4 * (size_t)max_slot,
                          hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    const uint64_t next = *reinterpret_cast<unsigned long long*>(S->hctrl + 8);
    const int was_busy = S->busy;
    S->busy = *reinterpret_cast<int*>(S->hctrl + 16);
    // every entry below min(next, tail) has been read by a slot
    const uint64_t consumed = std::min<uint64_t>(next, S->tail);
    {
        // the device's consumption rate while it had entries to read
        const auto now = Clock::now();
        const double dt = std::chrono::duration<double>(now - S->t_round).count();
        if (S->t_round != Clock::time_point{} && dt > 0 && consumed > S->consumed && S->busy > 0) {
            const double r = (double)(consumed - S->consumed) / dt;
            S->rate = S->rate > 0 ? 0.9 * S->rate + 0.1 * r : r;
        }
        S->t_round = now;
    }
    S->consumed = consumed;
    const uint32_t* done = reinterpret_cast<const uint32_t*>(S->hctrl + CTRL_DONE_OFF);
    std::vector<Request*> finished;
    for (Request* r : live)
        if (r->fed == r->npairs && done[r->slot] == r->npairs) finished.push_back(r);
    for (Request* r : finished) {
        rc = finish(S, r);
        if (rc) return rc;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (was_busy > 0) {  // busy slots as seen at the round's end, for each of its sweeps
            S->st_sweeps += (uint64_t)S->iters;
            S->st_busy += (uint64_t)was_busy * (uint64_t)S->iters;
        }
        const auto now = Clock::now();
        for (Request* r : finished) {
            r->state = RQ_DONE;
            r->t_done = now;
            release(S, r);
            S->live.erase(std::find(S->live.begin(), S->live.end(), r));
            S->st_requests++;
            S->st_pairs += r->npairs;
        }
    }
    if (!finished.empty()) S->cv_done.notify_all();
    return KESTREL_OK;
}

void fail_all(ServeState* S, const std::string& msg) {
    for (Request* r : S->live) {
        r->state = RQ_FAILED;
        r->err = msg;
    }
    for (Request* r : S->pending) {
        r->state = RQ_FAILED;
        r->err = msg;
    }
    S->live.clear();
    S->pending.clear();
}

void engine(ServeState* S) {
    if (hipSetDevice(S->c->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->failed = true;
        S->fail_msg = "serving engine: hipSetDevice failed";
        fail_all(S, S->fail_msg);
        S->cv_done.notify_all();
        return;
    }
    for (;;) {
        bool idle = false;
        const int rc = engine_round(S, &idle);
        if (rc) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->failed = true;
            S->fail_msg = "serving engine: " + last_error_message();
            fail_all(S, S->fail_msg);
            S->cv_done.notify_all();
            return;
        }
        if (idle) return;  // stop requested
    }
}

int report(ServeState* S, std::unordered_map<uint64_t, std::unique_ptr<Request>>::iterator it, int* done,
           double* latency_ms) {
    Request* r = it->second.get();
    if (r->state == RQ_FAILED) {
        const std::string msg = r->err;
        S->reqs.erase(it);
        return fail(KESTREL_EDEVICE, "%s", msg.c_str());
    }
    if (r->state != RQ_DONE) {
        if (done) *done = 0;
        return KESTREL_OK;
    }
    if (done) *done = 1;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    S->reqs.erase(it);
    return KESTREL_OK;
}

}  // namespace

void serve_release(ServeState* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_engine.notify_all();
    // the engine leaves at the start of its next round; only then are its
    // requests failed, so it never finds them gone mid-round
    if (S->th.joinable()) S->th.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        fail_all(S, "serving stopped");
    }
    S->cv_done.notify_all();
    delete S;
}

extern "C" {

int kestrel_serve_start(kestrel_cohort* c, const kestrel_tune* tune, uint32_t max_query_rows, uint64_t max_pairs) {
    if (!c || !tune) return fail(KESTREL_EINVAL, "null argument");
    if (c->gt) return fail(KESTREL_EINVAL, "serving fits the ML estimator; this cohort was created for -f GT");
    if (c->soft) return fail(KESTREL_EINVAL, "serving takes 2-bit hard calls; this cohort holds likelihood rows");
    if (!c->parts.empty()) return fail(KESTREL_EINVAL, "serving takes biallelic sites; this cohort has 3- or 4-allele sites");
    if (c->serve) return fail(KESTREL_EINVAL, "serving is already started on this cohort");
    if (c->n_kept == 0) return fail(KESTREL_EINVAL, "the cohort has no kept site");
    if (max_query_rows == 0 || max_pairs == 0) return fail(KESTREL_EINVAL, "max_query_rows and max_pairs must be > 0");
    if (max_pairs > 0xffffffffull) return fail(KESTREL_EINVAL, "max_pairs must fit 32 bits");
    if ((uint64_t)c->n_samples + max_query_rows > 0xffffffffull) return fail(KESTREL_EINVAL, "too many rows");
    if (tune->sqp_max_iter == 0 || tune->qp_max_iter == 0 || tune->bls_max_iter == 0)
        return fail(KESTREL_EINVAL, "iteration limits must be non-zero");
    if (!eps_in_domain(tune->epsilon))
        return fail(KESTREL_EINVAL, "epsilon %g outside [1e-18, 1e14] for a pair fit", tune->epsilon);
    HIPCHK(hipSetDevice(c->device));
    std::unique_ptr<ServeState> S(new ServeState());
    S->c = c;
    S->tune = *tune;
    S->dt = to_dev(*tune);
    if (c->mixed) S->dt.recond_margin = RECOND_MARGIN;
    S->qmax = max_query_rows;
    S->n_cohort = c->n_samples;
    S->row_words = c->row_words;
    S->out_cap = max_pairs;
    S->policy = (int)c->opt_serve_policy;
    S->rb = (c->n_sites + 3) / 4;
    HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));

    // slots as kestrel_fit_pairs: the resident sweep blocks over the segments,
    // per group; two groups on two streams unless option "streams" is 1
    // (the class-sorted form: one pair a wavefront, sixteen a CU)
    const uint64_t resident = (uint64_t)std::max(1, c->n_cu) * 2;
    S->nseg = c->nseg;
    S->cls = c->class_form();
    const uint64_t slot_blocks = std::max<uint64_t>(1, resident / S->nseg);
    uint64_t ns = c->opt_slots > 0 ? (uint64_t)c->opt_slots
                  : S->cls        ? (uint64_t)std::max(1, c->n_cu) * 16
                                  : slot_blocks * SWEEP_TPB;
    ns = (ns + SWEEP_TPB - 1) / SWEEP_TPB * SWEEP_TPB;
    S->ngroups = c->opt_streams >= 2 ? 2 : 1;
    if (S->cls) {  // the slots' lists within the device memory they may take
        const uint64_t cap = class_slot_cap(c->class_blocks(), S->ngroups, c->opt_class_list_mib);
        if (cap == 0)
            return fail(KESTREL_EDEVICE, "device memory for the class-sorted sweep's lists of %d x 256 slots is not "
                        "free (option class_sweep 0 selects the site-table sweep)", S->ngroups);
        ns = std::min(ns, cap);
    }
    S->ns = (uint32_t)ns;
    // a backlog of a quarter of the slots per sweep of a round refills
    // finished slots every step while keeping a newly admitted request close
    // to the front
    S->backlog = std::max<uint64_t>(FEED_CHUNK, ns / 4 * S->ngroups * S->iters);
    S->ring_cap = 2 * S->backlog;

    const size_t rows = (size_t)c->n_samples + max_query_rows;
    HIPCHK(S->geno.alloc(rows * S->row_words));
    HIPCHK(hipMemcpyAsync(S->geno.p, c->geno.p, (size_t)c->n_samples * S->row_words * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, S->stream));
    std::vector<uint32_t> kept;
    for (uint32_t s = 0; s < c->n_sites; s++)
        if (c->kept_mask[s]) kept.push_back(s);
    HIPCHK(S->kept.alloc(kept.size()));
    HIPCHK(hipMemcpyAsync(S->kept.p, kept.data(), kept.size() * sizeof(uint32_t), hipMemcpyHostToDevice, S->stream));
    HIPCHK(S->seg.alloc(S->nseg + 1));
    HIPCHK(hipMemcpyAsync(S->seg.p, c->seg.p, (S->nseg + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, S->stream));
    HIPCHK(S->qstage.alloc((size_t)max_query_rows * S->rb));
    HIPCHK(S->qrows.alloc(max_query_rows));
    HIPCHK(S->ring.alloc(4 * S->ring_cap));
    HIPCHK(S->outpool.alloc(max_pairs));
    HIPCHK(S->ctrl.alloc(CTRL_DONE_OFF + 4 * (size_t)MAX_REQ));
    HIPCHK(hipMemsetAsync(S->ctrl.p, 0, S->ctrl.n, S->stream));
    if (S->ngroups == 2) {
        HIPCHK(hipStreamCreateWithFlags(&S->stream2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&S->fed, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&S->swept, hipEventDisableTiming));
    }
    for (int g = 0; g < S->ngroups; g++) {
        ServeState::Group& G = S->grp[g];
        G.stream = g == 0 ? S->stream : S->stream2;
        HIPCHK(G.slots.alloc(ns));
        HIPCHK(G.sx.alloc(ns));
        HIPCHK(G.sy.alloc(ns));
        HIPCHK(G.pt.alloc(9 * ns));
        HIPCHK(G.partial.alloc((size_t)S->nseg * NPART * ns));
        HIPCHK(G.active.alloc(ns));
        HIPCHK(hipMemsetAsync(G.slots.p, 0, ns * sizeof(SlotState), S->stream));  // PH_IDLE
        HIPCHK(hipMemsetAsync(G.sx.p, 0, ns * sizeof(uint32_t), S->stream));
        HIPCHK(hipMemsetAsync(G.sy.p, 0, ns * sizeof(uint32_t), S->stream));
        HIPCHK(hipMemsetAsync(G.active.p, 0, ns * sizeof(int), S->stream));
        if (S->cls) {
            const int rc = G.cls.alloc((uint32_t)ns, c->class_blocks(), S->stream);
            if (rc) return rc;
        }
    }
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&S->hctrl), S->ctrl.n, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&S->hent), S->backlog * 16, hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(S->stream));
    S->device_bytes = rows * S->row_words * 4 + kept.size() * 4 + (S->nseg + 1) * 4 + S->qstage.n + 4 * S->qrows.n +
                      16 * S->ring_cap + max_pairs * sizeof(kestrel_output) + S->ctrl.n +
                      S->ngroups * (ns * sizeof(SlotState) + 8 * ns + 72 * ns + (size_t)S->nseg * NPART * ns * 8 + 4 * ns +
                                    (S->cls ? ns * ((size_t)c->class_blocks() * (CLASS_STRIDE + 8) * 2 + 8) : 0));

    for (uint32_t i = 0; i < MAX_REQ; i++) S->free_slots.push_back(MAX_REQ - 1 - i);
    for (uint32_t i = 0; i < max_query_rows; i++) S->free_rows.push_back(c->n_samples + max_query_rows - 1 - i);
    S->outs.init(max_pairs);
    ServeState* raw = S.release();
    raw->th = std::thread(engine, raw);
    c->serve = raw;
    return KESTREL_OK;
}

int kestrel_submit_queries(kestrel_cohort* c, const uint8_t* query_packed, uint32_t n_query, size_t row_stride_bytes,
                           const uint32_t* cohort_subset, size_t n_subset, kestrel_output* out,
                           uint64_t* request_id) {
    if (!c || !query_packed || !out || !request_id) return fail(KESTREL_EINVAL, "null argument");
    ServeState* S = c->serve;
    if (!S) return fail(KESTREL_EINVAL, "serving is not started on this cohort");
    if (n_query == 0) return fail(KESTREL_EINVAL, "a request needs at least one query sample");
    if (n_query > S->qmax)
        return fail(KESTREL_EINVAL, "request of %u query samples exceeds max_query_rows %u", n_query, S->qmax);
    if (row_stride_bytes < S->rb) return fail(KESTREL_EINVAL, "row stride too small");
    const size_t n_sub = cohort_subset ? n_subset : c->n_samples;
    if (n_sub == 0) return fail(KESTREL_EINVAL, "empty cohort subset");
    if (cohort_subset)
        for (size_t j = 0; j < n_subset; j++)
            if (cohort_subset[j] >= c->n_samples)
                return fail(KESTREL_EINVAL, "subset entry %zu names sample %u of %u", j, cohort_subset[j],
                            c->n_samples);
    const uint64_t npairs = (uint64_t)n_query * n_sub;
    if (npairs > S->out_cap)
        return fail(KESTREL_EINVAL, "request of %llu pairs exceeds max_pairs %llu", (unsigned long long)npairs,
                    (unsigned long long)S->out_cap);
    std::unique_ptr<Request> r(new Request());
    r->n_query = n_query;
    r->n_sub = (uint32_t)n_sub;
    r->npairs = npairs;
    r->out = out;
    r->qdata.resize((size_t)n_query * S->rb);
    for (uint32_t i = 0; i < n_query; i++)
        std::memcpy(r->qdata.data() + (size_t)i * S->rb, query_packed + (size_t)i * row_stride_bytes, S->rb);
    if (cohort_subset) r->subset.assign(cohort_subset, cohort_subset + n_subset);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->failed) return fail(KESTREL_EDEVICE, "%s", S->fail_msg.c_str());
        r->id = S->next_id++;
        r->t_submit = Clock::now();
        *request_id = r->id;
        S->pending.push_back(r.get());
        S->reqs.emplace(r->id, std::move(r));
    }
    S->cv_engine.notify_all();
    return KESTREL_OK;
}

int kestrel_poll(kestrel_cohort* c, uint64_t request_id, int* done, double* latency_ms) {
    if (!c || !done) return fail(KESTREL_EINVAL, "null argument");
    ServeState* S = c->serve;
    if (!S) return fail(KESTREL_EINVAL, "serving is not started on this cohort");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end()) return fail(KESTREL_EINVAL, "unknown request %llu", (unsigned long long)request_id);
    return report(S, it, done, latency_ms);
}

int kestrel_wait(kestrel_cohort* c, uint64_t request_id, double* latency_ms) {
    if (!c) return fail(KESTREL_EINVAL, "null argument");
    ServeState* S = c->serve;
    if (!S) return fail(KESTREL_EINVAL, "serving is not started on this cohort");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end()) return fail(KESTREL_EINVAL, "unknown request %llu", (unsigned long long)request_id);
    Request* r = it->second.get();
    S->cv_done.wait(lk, [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; });
    int done = 0;
    return report(S, it, &done, latency_ms);
}

int kestrel_serve_get_stats(kestrel_cohort* c, kestrel_serve_stats* st) {
    if (!c || !st) return fail(KESTREL_EINVAL, "null argument");
    ServeState* S = c->serve;
    if (!S) return fail(KESTREL_EINVAL, "serving is not started on this cohort");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof *st);
    st->requests_done = S->st_requests;
    st->pairs_done = S->st_pairs;
    st->sweeps = S->st_sweeps;
    st->busy_slot_sweeps = S->st_busy;
    st->device_bytes = S->device_bytes;
    st->slots = S->ns * (uint32_t)S->ngroups;
    st->live = (uint32_t)S->live.size();
    st->pending = (uint32_t)S->pending.size();
    return KESTREL_OK;
}

int kestrel_serve_stop(kestrel_cohort* c) {
    if (!c) return fail(KESTREL_EINVAL, "null argument");
    ServeState* S = c->serve;
    c->serve = nullptr;
    serve_release(S);
    return KESTREL_OK;
}

}  // extern "C"
