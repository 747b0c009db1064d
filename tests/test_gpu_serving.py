"""GPU: the server's results equal a plan's, alone and in a mix of concurrent requests; admission,
rejection and cancellation."""
import sys
import time

import numpy as np
import pytest

import gsofatpds_amd as g

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def digest_arrays(nzu, flags):
    return g.fnv1a_array(nzu.astype(np.int32)), g.fnv1a_array(flags.astype(np.uint8))


def plan_reference(p, **opt):
    with g.Plan(p, **opt) as plan:
        plan.run()
        r = plan.result()
        return r.as_dict(), digest_arrays(plan.nzu(), plan.supernode_flags())


KEYS = ("fill_count", "nnz_F", "nnz_L", "nnz_U", "edge_checks", "reach", "supernodes")


def test_server_alone_and_in_the_mix():
    members = [6, 11, 2, 8, 9, 12, 14, 16]
    pats = [g.generate_workload(g.WL_STREAM, m) if m % 2 == 0 else g.generate_workload(g.WL_BATCH256, m)
            for m in members]
    pats = [p for p in pats if p.n <= 60000][:6]
    say(f"[serve] n = {[p.n for p in pats]}")
    refs = [plan_reference(p) for p in pats]
    srv = g.Server(slots=4)
    try:
        # alone: one at a time
        for p, (ref, dig) in zip(pats, refs):
            rid = srv.submit(p)
            r = srv.wait(rid, 120.0).as_dict()
            info = srv.info(rid)
            assert all(r[k] == ref[k] for k in KEYS)
            assert (info["nzu_digest"], info["supernode_digest"]) == dig
        # in the mix: all at once, twice over, so requests queue and overlap
        ids = [srv.submit(p) for p in pats + pats]
        for k, rid in enumerate(ids):
            r = srv.wait(rid, 300.0).as_dict()
            ref, dig = refs[k % len(pats)]
            assert all(r[key] == ref[key] for key in KEYS), (k, r, ref)
            info = srv.info(rid)
            assert (info["nzu_digest"], info["supernode_digest"]) == dig
        st = srv.stats()
        say(f"[serve] stats {st}")
        assert st["peak_running"] >= 2  # requests really overlapped
    finally:
        srv.stop()


def test_server_groups_queued_requests():
    """One slot: requests that queue behind a running one are taken as one group (one launch
    sequence); each result equals the plan's, grouped or not."""
    members = [6, 11, 2, 8, 9, 12]
    pats = [g.generate_workload(g.WL_STREAM, m) if m % 2 == 0 else g.generate_workload(g.WL_BATCH256, m)
            for m in members]
    pats = [g.generate_workload(g.WL_GRID2D_659033)] + [p for p in pats if p.n <= 60000]  # the first one runs long
    refs = [plan_reference(p) for p in pats]
    for max_group in (1, 16):
        srv = g.Server(slots=1, max_group=max_group)
        try:
            order = list(range(len(pats))) + list(range(1, len(pats)))
            ids = [srv.submit(pats[i]) for i in order]
            for i, rid in zip(order, ids):
                r = srv.wait(rid, 300.0).as_dict()
                ref, dig = refs[i]
                assert all(r[key] == ref[key] for key in KEYS), (max_group, i)
                info = srv.info(rid)
                assert (info["nzu_digest"], info["supernode_digest"]) == dig
            st = srv.stats()
            say(f"[serve] max_group {max_group}: {st}")
            if max_group == 1:
                assert st["peak_running"] == 1
            else:
                assert st["peak_running"] >= 2  # a group formed on the single slot
        finally:
            srv.stop()


def test_server_rejects_what_never_fits_and_cancels():
    small = g.generate_workload(g.WL_GRID2D_4096)
    srv = g.Server(slots=1, hbm_budget_bytes=64 << 20)
    try:
        big = g.generate_workload(g.WL_GRID3D_110592)
        rid = srv.submit(big)
        with pytest.raises(g.GsofaError) as e:
            srv.wait(rid, 10.0)
        assert e.value.code == g.ERR_REJECTED
        # queue three; cancel the last while it is still pending
        a = srv.submit(small)
        b = srv.submit(small)
        c = srv.submit(small)
        srv.cancel(c)
        ra, rb = srv.wait(a, 60.0), srv.wait(b, 60.0)
        assert ra.fill_count == rb.fill_count
        with pytest.raises(g.GsofaError) as e:
            srv.wait(c, 10.0)
        assert e.value.code == g.ERR_CANCELLED
    finally:
        srv.stop()


def test_poisson_stream_alone_and_in_the_mix():
    """96 requests of the served stream (n log-uniform in [1,000, 500,000], ND stencils and
    Erdos-Renyi patterns) arriving as a seeded Poisson process at 200 per second, so that they
    arrive while others run: each result (totals, nzU and supernode digests) equals the same
    request run alone as a plan. Released requests leave nothing behind."""
    from concurrent.futures import ThreadPoolExecutor
    nreq = 96
    with ThreadPoolExecutor(8) as ex:
        pats = list(ex.map(lambda k: g.generate_workload(g.WL_STREAM, k), range(nreq)))
    ns = [p.n for p in pats]
    say(f"[stream] {nreq} requests, n from {min(ns)} to {max(ns)}")
    assert max(ns) > 250000 and min(ns) < 5000
    refs = [plan_reference(p) for p in pats]
    gaps = np.random.default_rng(0x5EED0005).exponential(1.0 / 200.0, nreq)
    srv = g.Server(slots=4)
    try:
        ids = []
        t = time.perf_counter()
        for p, gap in zip(pats, gaps):
            t += gap
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids.append(srv.submit(p))
        for k, rid in enumerate(ids):
            r = srv.wait(rid, 300.0).as_dict()
            ref, dig = refs[k]
            assert all(r[key] == ref[key] for key in KEYS), (k, pats[k].n)
            info = srv.info(rid)
            assert (info["nzu_digest"], info["supernode_digest"]) == dig, k
        st = srv.stats()
        say(f"[stream] stats {st}")
        assert st["peak_running"] >= 2 and st["completed"] == nreq
        before = srv.retained()
        assert before["requests"] == nreq
        for rid in ids:
            srv.release(rid)
        after = srv.retained()
        say(f"[stream] retained before release {before}, after {after}")
        assert after == {"requests": 0, "host_bytes": 0}
        with pytest.raises(g.GsofaError):
            srv.info(ids[0])  # released ids are unknown
    finally:
        srv.stop()


def test_server_cancels_a_running_request_and_bounds_memory():
    """A request cancelled while it runs on the device reports CANCELLED and the server stays
    usable; finished requests keep only their results (no pattern), so retained host bytes stay
    small until they are released."""
    big = g.generate_workload(g.WL_GRID2D_659033)
    small = g.generate_workload(g.WL_GRID2D_4096)
    srv = g.Server(slots=1, max_group=1)
    try:
        rid = srv.submit(big, bitmap_in_hbm=1)  # about 30 ms on the device
        t_end = time.time() + 30
        while srv.stats()["running"] == 0 and time.time() < t_end:
            time.sleep(0.0005)
        assert srv.poll(rid) is None  # still running
        srv.cancel(rid)
        with pytest.raises(g.GsofaError) as e:
            srv.wait(rid, 60.0)
        assert e.value.code == g.ERR_CANCELLED
        with pytest.raises(g.GsofaError) as e:
            srv.release(srv.submit(big))  # queued or running: not releasable
        assert e.value.code == g.ERR_STATE
        ids = [srv.submit(small) for _ in range(20)]
        ref = plan_reference(small)[0]
        for rid in ids:
            r = srv.wait(rid, 60.0).as_dict()
            assert all(r[k] == ref[k] for k in KEYS)
        kept = srv.retained()
        say(f"[serve] retained {kept}")
        assert kept["requests"] == 22 and kept["host_bytes"] < 22 * 4096  # no patterns kept
        with pytest.raises(g.GsofaError) as e:
            srv.submit(small, structure=1)
        assert e.value.code == g.ERR_INVALID_ARG
    finally:
        srv.stop()
