"""The request loop (sd_server): each result is the same alone and in the mix,
small requests overtake large ones, admission honours the memory budget.

kNN results are compared bit for bit with the CPU oracle and with one-shot
sparseKNN(X, Y) calls.  The SNN graph of a request is checked against a brute
force over all query pairs: queries i and j are joined with weight
|S_i & S_j| / |S_i | S_j| when that exceeds prune, where S_i is the set of
i's neighbours in Y.
"""
import os
import threading
import time

import numpy as np
import pytest
import scipy.sparse as sp

import sparsedist_amd as sd
from sparsedist_amd import _lib, synth
from conftest import bits

pytestmark = pytest.mark.gpu

PRUNE = 1 / 15
NT = min(16, os.cpu_count() or 1)


def same_as_oracle(oracle_lib, res, X, Y, k, method, cols=None):
    """Served rows (all, or the listed query columns) equal the oracle's."""
    if cols is None:
        cols = np.arange(X.shape[1])
    oi, od, _ = oracle_lib.knn(X[:, cols], Y, k=k, method=method, nthreads=NT)
    assert np.array_equal(res.idx[cols], oi), (method, k)
    assert np.array_equal(bits(res.dist[cols]), bits(od)), (method, k)


def brute_shared_snn(idx, prune):
    n, k = idx.shape
    sets = [set(int(v) for v in idx[i] if v != sd.NA_INTEGER) for i in range(n)]
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            u = len(sets[i] & sets[j])
            if u == 0:
                continue
            w = u / (len(sets[i]) + len(sets[j]) - u)
            if w > prune:
                rows += [i, j]
                cols += [j, i]
                vals += [w, w]
    return sp.csc_matrix((vals, (rows, cols)), shape=(n, n))


def same_graph(a, b):
    a = a.tocsc()
    b = b.tocsc()
    a.sort_indices()
    b.sort_indices()
    assert a.shape == b.shape
    assert np.array_equal(a.indptr, b.indptr)
    assert np.array_equal(a.indices, b.indices)
    assert np.array_equal(bits(a.data), bits(b.data))


@pytest.fixture(scope="module")
def ref_and_requests():
    Y = synth.nb_log1p(400, 6000, 0.05, 71)
    sizes = [1, 3, 17, 250, 900, 2500, 40, 7, 1200, 64, 333, 2]
    reqs = []
    for t, n in enumerate(sizes):
        X = synth.nb_log1p(400, n, 0.05, 100 + t)
        reqs.append(dict(X=X, k=(10, 20, 30)[t % 3], method=("cosine", "binary")[t % 2],
                         snn=(t % 2 == 0) and n <= 900))
    return Y, reqs


@pytest.mark.parametrize("graphs", [False, True])
def test_server_same_alone_and_in_mix(ref_and_requests, oracle_lib, gpu_ctx, graphs):
    Y, reqs = ref_and_requests
    dY = sd.DeviceMatrix(Y)
    with sd.Server(dY, slots=4, chunk_cols=256, graphs=graphs) as srv:
        tickets = [srv.submit(r["X"], k=r["k"], method=r["method"], snn=r["snn"], prune=PRUNE)
                   for r in reqs]
        mixed = [srv.result(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["pending"] == 0 and st["active"] == 0
    assert st["reserved_bytes"] == 0
    for r, m in zip(reqs, mixed):
        same_as_oracle(oracle_lib, m, r["X"], Y, r["k"], r["method"])
        # one-shot call on the user's context
        alone = sd.sparseKNN(r["X"], dY, k=r["k"], method=r["method"], verbose=False)
        assert np.array_equal(m.idx, alone.idx)
        assert np.array_equal(bits(m.dist), bits(alone.dist))
        if r["snn"]:
            same_graph(m["snn"], brute_shared_snn(m.idx, PRUNE))
    # the same requests alone on a one-slot server, one work item each
    with sd.Server(dY, slots=1, chunk_cols=100000) as srv1:
        for r, m in zip(reqs, mixed):
            if not r["snn"]:
                continue
            a = srv1.result(srv1.submit(r["X"], k=r["k"], method=r["method"], snn=True,
                                        prune=PRUNE))
            assert np.array_equal(a.idx, m.idx)
            same_graph(a["snn"], m["snn"])


def test_small_request_overtakes_large(oracle_lib, gpu_ctx):
    """A large request (47 work items) and a small one submitted after it to
    a paused one-slot server: once resumed, the slot takes the request with
    the fewest unscheduled columns first, so the small one finishes first."""
    Y = synth.nb_log1p(300, 5000, 0.05, 72)
    big = synth.nb_log1p(300, 6000, 0.05, 73)
    small = synth.nb_log1p(300, 5, 0.05, 74)
    with sd.Server(Y, slots=1, chunk_cols=128, paused=True) as srv:
        tb = srv.submit(big, k=10, method="cosine")
        ts = srv.submit(small, k=10, method="cosine")
        srv.resume()
        rb = srv.result(tb)
        rs = srv.result(ts)
    assert rs["info"]["finish_order"] == 1 and rb["info"]["finish_order"] == 2
    same_as_oracle(oracle_lib, rs, small, Y, 10, "cosine")
    same_as_oracle(oracle_lib, rb, big, Y, 10, "cosine", cols=np.arange(0, 6000, 37))


def test_admission_lets_fitting_requests_pass(oracle_lib, gpu_ctx):
    """A request whose reservation does not fit yet does not hold back a later
    one that fits (it is admitted once the budget frees, here when the server
    resumes and the first request completes)."""
    Y = synth.nb_log1p(200, 3000, 0.05, 86)
    first = synth.nb_log1p(200, 30, 0.05, 87)
    big = synth.nb_log1p(200, 3000, 0.05, 88)
    tiny = synth.nb_log1p(200, 3, 0.05, 89)
    reserve = {}
    with sd.Server(Y, slots=1, paused=True) as probe:  # each request's reservation
        for name, X, snn in (("first", first, False), ("big", big, True), ("tiny", tiny, False)):
            before = probe.stats()["reserved_bytes"]
            probe.submit(X, k=20, method="binary", snn=snn, prune=PRUNE)
            reserve[name] = probe.stats()["reserved_bytes"] - before
    # the first and the tiny request fit together, the big one fits alone but
    # not beside the first
    budget = max(reserve["big"], reserve["first"] + reserve["tiny"])
    assert reserve["first"] + reserve["big"] > budget
    with sd.Server(Y, slots=1, budget_bytes=budget, paused=True) as srv:
        t1 = srv.submit(first, k=20, method="binary")
        tb = srv.submit(big, k=20, method="binary", snn=True, prune=PRUNE)
        tt = srv.submit(tiny, k=20, method="binary")
        st = srv.stats()
        assert st["admitted"] == 2 and st["pending"] == 1, st
        srv.resume()
        r1, rb, rt = srv.result(t1), srv.result(tb), srv.result(tt)
    same_as_oracle(oracle_lib, r1, first, Y, 20, "binary")
    same_as_oracle(oracle_lib, rt, tiny, Y, 20, "binary")
    same_as_oracle(oracle_lib, rb, big, Y, 20, "binary", cols=np.arange(0, 3000, 29))
    assert rb["snn"].shape == (3000, 3000)


def test_small_lane_admits_behind_a_full_budget(oracle_lib, gpu_ctx):
    """A small request that the budget cannot take (a large one holds it and
    another large one waits) is admitted into the small-request lane, a
    sixteenth of the budget beyond it, instead of waiting for the large
    ones; the waiting large request is admitted once the budget frees."""
    Y = synth.nb_log1p(200, 3000, 0.05, 95)
    big1 = synth.nb_log1p(200, 3000, 0.05, 96)
    big2 = synth.nb_log1p(200, 3000, 0.05, 97)
    tiny = synth.nb_log1p(200, 3, 0.05, 98)
    reserve = {}
    with sd.Server(Y, slots=1, paused=True) as probe:
        for name, X in (("big", big1), ("tiny", tiny)):
            before = probe.stats()["reserved_bytes"]
            probe.submit(X, k=20, method="cosine", snn=True, prune=PRUNE)
            reserve[name] = probe.stats()["reserved_bytes"] - before
    budget = max(reserve["big"], 16 * reserve["tiny"])
    assert 2 * reserve["big"] > budget  # the second large request waits
    with sd.Server(Y, slots=1, budget_bytes=budget, paused=True) as srv:
        t1 = srv.submit(big1, k=20, method="cosine", snn=True, prune=PRUNE)
        t2 = srv.submit(big2, k=20, method="cosine", snn=True, prune=PRUNE)
        tt = srv.submit(tiny, k=20, method="cosine", snn=True, prune=PRUNE)
        st = srv.stats()
        assert st["admitted"] == 2 and st["pending"] == 1 and st["lane_admitted"] == 1, st
        srv.resume()
        r1, r2, rt = srv.result(t1), srv.result(t2), srv.result(tt)
        st = srv.stats()
    assert st["completed"] == 3 and st["reserved_bytes"] == 0, st
    same_as_oracle(oracle_lib, rt, tiny, Y, 20, "cosine")
    for r, X in ((r1, big1), (r2, big2)):
        same_as_oracle(oracle_lib, r, X, Y, 20, "cosine", cols=np.arange(0, 3000, 37))


def test_destroy_while_a_caller_waits(gpu_ctx):
    """sd_server_destroy while another thread is blocked in sd_server_wait:
    the waiter returns (its request cancelled) before the server is freed."""
    Y = synth.nb_log1p(200, 1000, 0.05, 93)
    X = synth.nb_log1p(200, 50, 0.05, 94)
    srv = sd.Server(Y, slots=1, paused=True)
    t = srv.submit(X, k=5, method="cosine")
    out = {}

    def waiter():
        out["info"] = srv.wait(t)

    th = threading.Thread(target=waiter)
    th.start()
    deadline = time.time() + 30
    while srv.stats()["waiting_callers"] < 1:
        assert time.time() < deadline, "the waiter never blocked"
        time.sleep(0.01)
    srv.close()
    th.join(timeout=30)
    assert not th.is_alive()
    assert out["info"]["done"] == 1 and out["info"]["status"] == _lib.SD_ECANCELLED


def test_close_with_work_in_flight(gpu_ctx):
    """Closing a server whose slots are busy cancels the work (polled between
    kernels) and returns."""
    Y = synth.nb_log1p(2000, 20000, 0.05, 95)
    srv = sd.Server(Y, slots=2, chunk_cols=512)
    ts = [srv.submit(synth.nb_log1p(2000, 4000, 0.05, 96 + i), k=10, method="cosine")
          for i in range(3)]
    time.sleep(0.2)
    t0 = time.perf_counter()
    srv.close()
    assert time.perf_counter() - t0 < 60
    del ts


def test_serving_workload_reference(oracle_lib, gpu_ctx):
    """The serving workload's reference rows and density (20000 rows, d =
    0.05, NB/log1p; 20000 columns here, so every work item runs the pre-pass
    and filtered pass): a mix of request sizes, methods and k served
    concurrently on four slots; sampled rows of every request equal the
    oracle's, and the SNN graphs of the small SNN requests equal the brute
    force."""
    Y = synth.nb_log1p(20000, 20000, 0.05, 531)
    dY = sd.DeviceMatrix(Y, gpu_ctx)
    sizes = [1, 6, 45, 300, 2100, 5000, 12, 700]
    reqs = [dict(X=synth.nb_log1p(20000, n, 0.05, 540 + t), k=(10, 20, 30)[t % 3],
                 method=("cosine", "binary")[t % 2], snn=t % 2 == 0) for t, n in enumerate(sizes)]
    with sd.Server(dY, slots=4, chunk_cols=4096, paused=True) as srv:
        ts = [srv.submit(r["X"], k=r["k"], method=r["method"], snn=r["snn"], prune=PRUNE)
              for r in reqs]
        srv.resume()
        res = [srv.result(t) for t in ts]
    rng = np.random.default_rng(532)
    for r, m in zip(reqs, res):
        n = r["X"].shape[1]
        cols = np.sort(rng.choice(n, min(n, 24), replace=False))
        same_as_oracle(oracle_lib, m, r["X"], Y, r["k"], r["method"], cols=cols)
        if r["snn"] and n <= 700:
            same_graph(m["snn"], brute_shared_snn(m.idx, PRUNE))


def test_budget_admission(gpu_ctx):
    Y = synth.nb_log1p(200, 3000, 0.05, 75)
    X = synth.nb_log1p(200, 50, 0.05, 76)
    with sd.Server(Y, slots=2, chunk_cols=64, budget_bytes=1 << 20) as srv:
        with pytest.raises(sd.SparseDistError) as e:
            srv.submit(X, k=5, method="binary")
        assert e.value.code == _lib.SD_EBUDGET
    # a budget that fits one request at a time (about 3.2 MB reserved each)
    # still serves all of them, in order
    with sd.Server(Y, slots=2, chunk_cols=64, budget_bytes=4 << 20) as srv:
        ts = [srv.submit(synth.nb_log1p(200, 40 + i, 0.05, 80 + i), k=5, method="binary")
              for i in range(4)]
        res = [srv.result(t) for t in ts]
        st = srv.stats()
    assert st["completed"] == 4 and st["admitted"] == 4
    for i, r in enumerate(res):
        alone = sd.sparseKNN(synth.nb_log1p(200, 40 + i, 0.05, 80 + i), Y, k=5,
                             method="binary", verbose=False)
        assert np.array_equal(r.idx, alone.idx)


def test_request_errors_are_reported(gpu_ctx):
    Y = synth.nb_log1p(200, 1000, 0.05, 77)
    with sd.Server(Y, slots=2) as srv:
        with pytest.raises(sd.SparseDistError) as e:
            srv.submit(synth.nb_log1p(150, 10, 0.05, 78), k=5)
        assert "same number of rows" in str(e.value)
        # non-finite cosine input fails inside the worker and is reported by result()
        X = synth.nb_log1p(200, 10, 0.05, 79).tolil()
        X[3, 2] = np.inf
        t = srv.submit(X.tocsc(), k=5, method="cosine")
        with pytest.raises(sd.SparseDistError) as e:
            srv.result(t)
        assert e.value.code == _lib.SD_ENONFINITE
        # the server keeps serving after a failed request
        ok = srv.result(srv.submit(synth.nb_log1p(200, 10, 0.05, 81), k=5, method="cosine"))
        assert ok.idx.shape == (10, 5)


def test_small_requests_share_a_tile(gpu_ctx):
    """Small requests with the same search are packed into one work item (a
    slab computes whole tiles of 256 query columns): while a large request
    holds the only slot, twenty small ones queue up, and the slot then serves
    them together; each result is still the one-shot search's, and a request
    with other parameters is never mixed in."""
    Y = synth.nb_log1p(300, 8000, 0.05, 91)
    big = synth.nb_log1p(300, 6000, 0.05, 92)
    small = [synth.nb_log1p(300, 3 + (i % 5), 0.05, 200 + i) for i in range(20)]
    other = synth.nb_log1p(300, 4, 0.05, 300)
    with sd.Server(Y, slots=1, chunk_cols=2048) as srv:
        tb = srv.submit(big, k=10, method="cosine")
        ts = [srv.submit(X, k=10, method="cosine") for X in small]
        to = srv.submit(other, k=15, method="cosine")
        res = [srv.result(t) for t in ts]
        ro = srv.result(to)
        srv.result(tb)
        st = srv.stats()
    assert st["packed_items"] >= 1
    for X, r in zip(small, res):
        alone = sd.sparseKNN(X, Y, k=10, method="cosine", verbose=False)
        assert np.array_equal(r.idx, alone.idx)
        assert np.array_equal(bits(r.dist), bits(alone.dist))
    alone = sd.sparseKNN(other, Y, k=15, method="cosine", verbose=False)
    assert np.array_equal(ro.idx, alone.idx)


def test_graph_items_replay_and_match_oracle(oracle_lib, gpu_ctx):
    """Work items of at most 256 columns replay a captured HIP graph.  The
    reference has 20000 columns (157 reference tiles), so the captured
    search is the two-pass one: threshold pre-pass, filtered pass, selection,
    re-rank.  One slot serves the requests one after another, so each graph
    is replayed several times, with ordinary items (more than 256 columns),
    one-shot searches on another context and host synchronisations in
    between; every served row equals the oracle's bit for bit."""
    Y = synth.nb_log1p(1000, 20000, 0.05, 81)
    dY = sd.DeviceMatrix(Y, gpu_ctx)
    plan = [(5, "cosine", 20), (256, "binary", 10), (40, "cosine", 20), (1, "binary", 10),
            (130, "cosine", 20), (300, "cosine", 20), (77, "binary", 10), (256, "cosine", 20),
            (12, "binary", 10), (200, "cosine", 20), (3, "cosine", 20), (90, "binary", 10)]
    with sd.Server(dY, slots=1, graphs=True) as srv:
        for t, (n, method, k) in enumerate(plan):
            X = synth.nb_log1p(1000, n, 0.05, 200 + t)
            res = srv.result(srv.submit(X, k=k, method=method))
            same_as_oracle(oracle_lib, res, X, Y, k, method)
            if t in (4, 8):  # eager work between replays
                one = sd.sparseKNN(X, dY, k=k, method=method, verbose=False)
                assert np.array_equal(one.idx, res.idx)
                assert np.array_equal(bits(one.dist), bits(res.dist))
        st = srv.stats()
    print("graph stats:", {f: st[f] for f in ("graph_items", "graph_captures", "graph_fallbacks")})
    assert st["graph_captures"] >= 2  # one per search parameters (cosine k=20, binary k=10)
    assert st["graph_items"] + st["graph_fallbacks"] >= 8  # 11 small items, 2 of them captured
    assert st["graph_items"] >= 5


def test_graph_items_off_gives_same_results(oracle_lib, gpu_ctx):
    """graphs=False (the default) serves every item as an ordinary search,
    with the same rows as the graph replays."""
    Y = synth.nb_log1p(1000, 20000, 0.05, 82)
    dY = sd.DeviceMatrix(Y, gpu_ctx)
    Xs = [synth.nb_log1p(1000, n, 0.05, 300 + n) for n in (9, 120, 9)]
    out = {}
    for graphs in (True, False):
        with sd.Server(dY, slots=1, graphs=graphs) as srv:
            out[graphs] = [srv.result(srv.submit(X, k=15, method="cosine")) for X in Xs]
            st = srv.stats()
        if not graphs:
            assert st["graph_items"] == 0 and st["graph_captures"] == 0
    for a, b, X in zip(out[True], out[False], Xs):
        assert np.array_equal(a.idx, b.idx)
        assert np.array_equal(bits(a.dist), bits(b.dist))
    same_as_oracle(oracle_lib, out[True][2], Xs[2], Y, 15, "cosine")
