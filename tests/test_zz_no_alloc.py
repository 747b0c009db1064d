"""The allocation-free entry point (gsofa_run_noalloc) against a plan on the same patterns, through
the C ABI: nzU, supernode flags and totals bit-exact, on calls in a row that reuse one scratch."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import gsofatpds_amd as g
import oracle
from conftest import GOLDEN


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def prepared(p):
    """Rows sorted with duplicates dropped: the row rule the entry point takes."""
    rows = [np.unique(np.asarray(p.adj[p.beg[v]:p.end[v]], np.int32)) for v in range(p.n)]
    rowptr = np.zeros(p.n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    adj = np.concatenate(rows).astype(np.int32) if rowptr[-1] else np.zeros(0, np.int32)
    return rowptr, adj


class DeviceBuffers:
    """Device memory through the HIP runtime the engine itself links (a framework's runtime in the
    same process would take the device away from it). The caller allocates; the entry point does not."""

    def __init__(self):
        g.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p.value)
        if a.nbytes:
            assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def run_noalloc(rowptr, adj, **opt):
    n = len(rowptr) - 1
    buf = DeviceBuffers()
    stream = g.Stream()
    try:
        d_rowptr = buf.upload(rowptr)
        d_adj = buf.upload(adj if len(adj) else np.zeros(1, np.int32))
        # sentinels: every output entry must be written
        nzu0 = np.full(n, -7, np.int32)
        flags0 = np.full(n, 0xAB, np.uint8)
        tot0 = np.full(len(g.NOALLOC_TOTALS), 0xFFFFFFFFFFFFFFFF, np.uint64)
        d_nzu, d_flags, d_tot = buf.upload(nzu0), buf.upload(flags0), buf.upload(tot0)
        g.run_noalloc(n, len(adj), d_rowptr, d_adj, d_nzu, d_flags, d_tot, stream=stream, **opt)
        g.device_synchronize()
        nzu, flags, t = buf.download(d_nzu, nzu0), buf.download(d_flags, flags0), buf.download(d_tot, tot0)
    finally:
        g.device_synchronize()
        buf.close()
        stream.close()
    return nzu, flags, dict(zip(g.NOALLOC_TOTALS, (int(x) for x in t)))


def check_against_plan(p, label, **opt):
    rowptr, adj = prepared(p)
    say(f"[noalloc] {label}: n={p.n} nnz={len(adj)} opt={opt}")
    nzu, flags, tot = run_noalloc(rowptr, adj, **opt)
    with g.Plan(p, structure=0, **opt) as plan:
        plan.run()
        r = plan.result()
        assert tot["input_error"] == 0
        for k in ("fill_count", "nnz_F", "nnz_L", "nnz_U", "edge_checks", "reach", "supernodes",
                  "supernodes_counted"):
            assert tot[k] == getattr(r, k), (k, tot[k], getattr(r, k))
        assert np.array_equal(nzu, plan.nzu())
        assert np.array_equal(flags, plan.supernode_flags())
    return nzu, flags, tot


@pytest.mark.gpu
def test_noalloc_equals_plan_on_calls_in_a_row():
    """Three calls on one scratch, each smaller than the one before, against a plan; the first two
    also against the oracle."""
    z = np.load(os.path.join(GOLDEN, "ds_11948.npz"))
    ds = g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])
    grid = g.generate(kind=g.GEN_STENCIL, dims=2, points=5, nx=256, ny=256, order=g.ORDER_ND, seed=3)
    er = g.generate(kind=g.GEN_ERDOS_RENYI, n_keep=20000, avg_degree_x100=400, seed=5)
    small = g.generate_workload(g.WL_GRID2D_4096)
    check_against_plan(grid, "grid 256x256 ND (65,536 rows, the row limit)")
    for p, label, opt in ((ds, "reference dataset", {}), (small, "grid2d_4096", {"chunk_size": 64})):
        nzu, flags, tot = check_against_plan(p, label, **opt)
        o = oracle.symbfact(p, chunk=opt.get("chunk_size", 128), structure=False)
        assert tot["fill_count"] == o["fill"]
        assert tot["supernodes_counted"] == o["nsup"]
        assert np.array_equal(nzu, o["nzu"])
        assert np.array_equal(flags, o["sup_flag"])
    check_against_plan(er, "random 20,000 rows, paired items", pair_directions=1)


@pytest.mark.gpu
def test_noalloc_reports_a_pattern_that_breaks_the_row_rule():
    """Unsorted rows and a column id outside [0, n) are caught on the device: input_error is set
    and the counts are those of the empty pattern (every row its own supernode start)."""
    n = 300
    rowptr = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    good = np.sort(np.stack([np.arange(n), (np.arange(n) + 1) % n], 1), axis=1).astype(np.int32).ravel()
    for label, bad in (("unsorted row", np.array([1, 0])), ("column id n", np.array([5, n]))):
        adj = good.copy()
        adj[10:12] = bad
        say(f"[noalloc] {label}")
        nzu, flags, tot = run_noalloc(rowptr, adj)
        assert tot["input_error"] == 1
        assert tot["fill_count"] == 0 and tot["nnz_F"] == n
        assert not nzu.any()
        assert tot["supernodes_counted"] == int(flags.sum())
    # and a valid call right after, on the same scratch
    p = g.Pattern.from_rowptr(rowptr.astype(np.int64), good)
    check_against_plan(p, "valid pattern after the rejected ones")


def test_noalloc_rejects_sizes_beyond_its_limits_on_the_host():
    """Size and option checks run before any device call (no GPU needed)."""
    fake = 1 << 20  # never dereferenced: the call fails before touching the device
    for n, nnz, opt in ((65537, 10, {}), (1000, (1 << 20) + 1, {}), (1000, 10, {"structure": 1}),
                        (1000, 10, {"autotune": 1}), (1000, 10, {"bitmap_in_hbm": 1})):
        with pytest.raises(g.GsofaError):
            g.run_noalloc(n, nnz, fake, fake, fake, fake, fake, **opt)
