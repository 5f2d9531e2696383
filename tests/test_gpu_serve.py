"""Serving engine (kestrel_serve_start / kestrel_submit_queries / kestrel_wait)
through the C ABI.

Bars: every request's records are bit-identical to the oracle fit of the same
(query, cohort sample) pairs in the device's summation order (the class-sorted
sweep's for FP64 cohorts, the default), bit-identical
alone and in a mix of concurrently admitted requests, and bit-identical to
kestrel_fit_pairs when a query repeats a cohort sample's row.
"""
import threading

import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, synth

pytestmark = pytest.mark.gpu

N_COHORT = 128


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(out, exp):
    assert len(out) == len(exp)
    keys = ["x", "y", "jacquard", "obj", "iters", "n_sites", "cond"]
    if "status" in exp.dtype.names:
        keys.append("status")
    for k in keys:
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} records differ"


@pytest.fixture(scope="module")
def served(O):
    c = synth.generate(N_COHORT + 40, 9000, seed=41, n_family_pairs=24, missing=0.02)
    cohort, queries = c.codes[:N_COHORT], c.codes[N_COHORT:]
    co = Cohort(cohort)
    af, kept = co.allele_freqs()
    co.serve_start(max_query_rows=48, max_pairs=40_000)
    yield co, cohort, queries, af, kept
    co.close()


def oracle_request(O, co, cohort, kept, af, qcodes, subset):
    """The oracle fit of a request: pairs (query, cohort sample) over the
    cohort's kept sites and allele frequencies, start point 1/9."""
    codes = np.vstack([cohort[:, kept], qcodes[:, kept]])
    sub = np.arange(N_COHORT) if subset is None else np.asarray(subset)
    q = np.repeat(np.arange(qcodes.shape[0]), len(sub))
    pairs = np.stack([N_COHORT + q, np.tile(sub, qcodes.shape[0])], 1)
    exp = O.fit_pairs(np.ascontiguousarray(codes), af[kept], pairs=pairs, order=O.ORDER_CLASS)
    exp["x"] = q
    return exp


def test_request_alone_matches_oracle(O, served):
    co, cohort, queries, af, kept = served
    say("case: serve 5 queries vs the whole cohort (640 pairs)")
    rid = co.submit(queries[:5])
    out, lat = co.wait(rid)
    assert lat > 0
    assert (out["status"] == 0).all()
    assert_bitexact(out, oracle_request(O, co, cohort, kept, af, queries[:5], None))


def test_query_repeating_a_cohort_row_matches_batch_fit(served):
    co, cohort, queries, af, kept = served
    say("case: query = cohort sample 7 vs subset, against kestrel_fit_pairs")
    sub = np.array([0, 3, 8, 9, 40, 41, 127, 100])  # sample 7 itself excluded: a self pair starts elsewhere
    out, _ = co.wait(co.submit(cohort[7:8], sub))
    exp = co.fit_pairs(np.stack([np.full(len(sub), 7), sub], 1))
    exp["x"] = 0
    assert_bitexact(out, exp)


def test_mix_equals_alone(O, served):
    """Requests admitted together (continuous batching, shared slots) give the
    same records as each request alone."""
    co, cohort, queries, af, kept = served
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(10):
        nq = int(rng.integers(1, 9))
        q0 = int(rng.integers(0, queries.shape[0] - nq + 1))
        sub = None if i % 3 == 0 else np.sort(rng.choice(N_COHORT, int(rng.integers(1, 90)), replace=False))
        reqs.append((queries[q0:q0 + nq], sub))
    say("case: 10 requests alone, one after the other")
    alone = [co.wait(co.submit(q, s))[0] for q, s in reqs]
    say("case: the same 10 requests submitted at once")
    ids = [co.submit(q, s) for q, s in reqs]
    mixed = [co.wait(i)[0] for i in ids]
    for a, m in zip(alone, mixed):
        assert_bitexact(m, a)
    say("case: oracle check of two mixed requests")
    for k in (1, 2):
        assert_bitexact(mixed[k], oracle_request(O, co, cohort, kept, af, *reqs[k]))


def test_admission_by_budget(served):
    """Requests larger than the free query rows wait for earlier ones; all
    finish and equal their results alone."""
    co, cohort, queries, af, kept = served
    say("case: 4 requests of 30 queries with 48 query rows")
    sub = np.arange(0, N_COHORT, 5)
    alone = co.wait(co.submit(queries[:30], sub))[0]
    ids = [co.submit(queries[:30], sub) for _ in range(4)]
    for i in ids:
        assert_bitexact(co.wait(i)[0], alone)


def test_submit_from_threads(served):
    co, cohort, queries, af, kept = served
    say("case: submit and wait from 4 threads")
    sub = np.arange(10, 60)
    alone = co.wait(co.submit(queries[3:6], sub))[0]
    res, errs = [], []

    def worker():
        try:
            for _ in range(3):
                res.append(co.wait(co.submit(queries[3:6], sub))[0])
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=worker) for _ in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs and len(res) == 12
    for r in res:
        assert_bitexact(r, alone)


def test_poll_and_stats(served):
    co, cohort, queries, af, kept = served
    say("case: poll and stats")
    before = co.serve_stats()
    rid = co.submit(queries[:2], np.arange(20))
    while True:
        done, lat, out = co.poll(rid)
        if done:
            break
    assert lat > 0 and len(out) == 40 and (out["x"] < 2).all()
    st = co.serve_stats()
    assert st["requests_done"] == before["requests_done"] + 1
    assert st["pairs_done"] == before["pairs_done"] + 40
    assert st["live"] == 0 and st["pending"] == 0 and st["slots"] > 0


def test_rejected_requests(served):
    co, cohort, queries, af, kept = served
    say("case: requests the engine rejects")
    with pytest.raises(KestrelError):
        co.submit(queries[:1], np.array([N_COHORT]))  # no such cohort sample
    with pytest.raises(KestrelError):
        co.submit(np.vstack([queries, queries])[:49])  # more than max_query_rows
    with pytest.raises(KestrelError):
        co.submit(queries[:40], np.arange(N_COHORT).repeat(8))  # more than max_pairs
    with pytest.raises(KestrelError):
        co.submit(queries[:1], np.array([], np.uint32))


def test_query_with_no_calls(served):
    co, cohort, queries, af, kept = served
    say("case: a query with no call")
    q = queries[:2].copy()
    q[1, :] = 3
    out, _ = co.wait(co.submit(q, np.arange(4)))
    assert (out["status"][:4] == 0).all()
    assert (out["status"][4:] == 1).all() and (out["n_sites"][4:] == 0).all()


def test_stop_fails_pending(served):
    co, cohort, queries, af, kept = served
    say("case: stop with requests queued, then restart")
    ids = [co.submit(queries[:30]) for _ in range(3)]
    co.serve_stop()
    for i in ids:
        with pytest.raises(KestrelError):
            co.wait(i)
    co.serve_start(max_query_rows=48, max_pairs=40_000)
    out, _ = co.wait(co.submit(queries[:1], np.arange(3)))
    assert (out["status"] == 0).all()


def test_round_robin_policy_same_results(served):
    co, cohort, queries, af, kept = served
    say("case: round-robin feed order, mix vs shortest-remaining-first alone")
    reqs = [(queries[0:6], None), (queries[6:7], np.arange(0, N_COHORT, 3)), (queries[7:12], np.arange(50))]
    alone = [co.wait(co.submit(q, s))[0] for q, s in reqs]
    co.serve_stop()
    co.set_option("serve_policy", 0)
    co.serve_start(max_query_rows=48, max_pairs=40_000)
    try:
        ids = [co.submit(q, s) for q, s in reqs]
        for i, a in zip(ids, alone):
            assert_bitexact(co.wait(i)[0], a)
    finally:
        co.serve_stop()
        co.set_option("serve_policy", 1)
        co.serve_start(max_query_rows=48, max_pairs=40_000)


def test_mixed_precision_cohort_serves_its_own_form():
    """A cohort made with the FP32-mixed form serves with that form: a query
    that repeats a cohort row gives kestrel_fit_pairs' records on the same
    cohort, bit for bit."""
    say("case: serving a KESTREL_MIXED cohort, query = cohort sample 5")
    c = synth.generate(96, 6000, seed=43, n_family_pairs=6, missing=0.02)
    with Cohort(c.codes, precision="mixed") as co:
        co.allele_freqs()
        co.serve_start(max_query_rows=8, max_pairs=4_000)
        sub = np.array([0, 1, 2, 17, 40, 95])
        out, _ = co.wait(co.submit(c.codes[5:6], sub))
        exp = co.fit_pairs(np.stack([np.full(len(sub), 5), sub], 1))
        exp["x"] = 0
        assert_bitexact(out, exp)
