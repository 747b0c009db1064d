"""MI355X symbolic-factorisation engine: Python view of the C ABI in ``include/gsofa.h``.

The engine is ``libgsofa.so`` next to this file (HIP kernels for gfx950 plus host C++). This
module only binds it with ctypes and converts arrays; every computation happens in the library.
If the library is missing, every entry point raises ``RuntimeError``: there is no fallback.

Interface mirror of the reference (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh):
  * ``symbfact_min_id(argv, rank)``   <- ``symbfact_min_id(args, argv, myrank, fill, time, nsup, teps)`` (:2529)
  * ``read_ascii(beg, end, csr)``      <- the reader inside ``symbfact_min_id`` (:2567-2616)
  * ``Plan`` / ``Batch``               structured forms of the same computation
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsofa.so")
CLI_PATH = os.path.join(_HERE, "gsofa_cli")
CONVERT_PATH = os.path.join(_HERE, "gsofa_convert")

# status codes (include/gsofa.h)
OK = 0
ERR_INVALID_ARG = 1
ERR_IO = 2
ERR_BAD_CSR = 3
ERR_HIP = 4
ERR_NO_DEVICE = 5
ERR_OUT_OF_MEMORY = 6
ERR_CAPACITY = 7
ERR_STATE = 8
ERR_REJECTED = 9
ERR_CANCELLED = 10
ERR_NOT_READY = 11

# workloads
WL_GRID2D_4096 = 1
WL_GRID3D_110592 = 2
WL_GRID2D_659033 = 3
WL_BATCH256 = 4
WL_STREAM = 5

GEN_STENCIL, GEN_ERDOS_RENYI = 0, 1
ORDER_NATURAL, ORDER_ND, ORDER_RCM = 0, 1, 2
DROP_ANY, DROP_LOWER, DROP_UPPER = 0, 1, 2


class GsofaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gsofa error {code}: {msg}")
        self.code = code


class CSR(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("beg", C.POINTER(C.c_int32)),
                ("end", C.POINTER(C.c_int32)), ("adj", C.POINTER(C.c_int32))]


class Options(C.Structure):
    _fields_ = [("chunk_size", C.c_int32), ("n_processes", C.c_int32), ("structure", C.c_int32),
                ("device", C.c_int32), ("heavy_threads", C.c_int32), ("heavy_wgs_per_cu", C.c_int32),
                ("light_wgs_per_cu", C.c_int32), ("queue_capacity", C.c_int32), ("chain_block", C.c_int32),
                ("pair_directions", C.c_int32), ("autotune", C.c_int32), ("bitmap_in_hbm", C.c_int32),
                ("split_blocks", C.c_int32), ("symmetric", C.c_int32), ("cu_share", C.c_int32),
                ("early_top", C.c_int32), ("pre_escalate", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("fill_count", C.c_uint64), ("nnz_F", C.c_uint64), ("nnz_L", C.c_uint64),
                ("nnz_U", C.c_uint64), ("edge_checks", C.c_uint64), ("reach", C.c_uint64),
                ("supernodes", C.c_int64), ("supernodes_counted", C.c_int64),
                ("heavy_sources", C.c_uint64), ("time_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib_handle = None

_SIGNATURES = {
    "gsofa_last_error": (C.c_char_p, []),
    "gsofa_generate_workload": (C.c_int, [C.c_int, C.c_int64, C.POINTER(C.c_void_p)]),
    "gsofa_generate": (C.c_int, [C.POINTER(C.c_int64), C.c_uint64, C.POINTER(C.c_void_p)]),
    "gsofa_host_matrix_csr": (C.c_int, [C.c_void_p, C.POINTER(CSR)]),
    "gsofa_host_matrix_free": (None, [C.c_void_p]),
    "gsofa_read_ascii": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "gsofa_write_ascii": (C.c_int, [C.POINTER(CSR), C.c_char_p, C.c_char_p, C.c_char_p]),
    "gsofa_plan_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "gsofa_batch_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "gsofa_read_binary": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "gsofa_write_binary": (C.c_int, [C.POINTER(CSR), C.c_char_p, C.c_char_p, C.c_int32, C.c_int32]),
    "gsofa_symbfact_min_id": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
    "gsofa_device_query": (C.c_int, [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.c_char_p, C.c_int32]),
    "gsofa_device_synchronize": (C.c_int, [C.c_int]),
    "gsofa_bandwidth_probe": (C.c_int, [C.c_int, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "gsofa_stream_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gsofa_stream_destroy": (C.c_int, [C.c_void_p]),
    "gsofa_options_default": (None, [C.POINTER(Options)]),
    "gsofa_plan_create": (C.c_int, [C.POINTER(CSR), C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "gsofa_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gsofa_sync": (C.c_int, [C.c_void_p]),
    "gsofa_result_get": (C.c_int, [C.c_void_p, C.POINTER(Result)]),
    "gsofa_nzu_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "gsofa_supernodes_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)]),
    "gsofa_structure_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gsofa_structure_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "gsofa_stage_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "gsofa_work_split": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    "gsofa_executed_work": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "gsofa_early_top": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "gsofa_plan_stage_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "gsofa_plan_destroy": (None, [C.c_void_p]),
    "gsofa_read_matrix_market": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "gsofa_read_permutation": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "gsofa_permute": (C.c_int, [C.POINTER(CSR), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_void_p)]),
    "gsofa_supernode_partition": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]),
    "gsofa_batch_supernode_partition": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gsofa_supernodal_structure_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                                  C.POINTER(C.c_int64)]),
    "gsofa_supernodal_structure_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                                 C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "gsofa_supernodal_from_structure": (C.c_int, [C.c_int32, C.POINTER(C.c_uint8), C.POINTER(C.c_int64),
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                                  C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "gsofa_batch_create": (C.c_int, [C.POINTER(CSR), C.c_int32, C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "gsofa_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gsofa_batch_sync": (C.c_int, [C.c_void_p]),
    "gsofa_batch_result_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Result)]),
    "gsofa_batch_nzu_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "gsofa_batch_supernodes_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint8)]),
    "gsofa_batch_destroy": (None, [C.c_void_p]),
    "gsofa_server_options_default": (None, [C.c_void_p]),
    "gsofa_server_start": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "gsofa_server_submit": (C.c_int, [C.c_void_p, C.POINTER(CSR), C.POINTER(Options), C.POINTER(C.c_int64)]),
    "gsofa_server_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Result)]),
    "gsofa_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(Result)]),
    "gsofa_server_cancel": (C.c_int, [C.c_void_p, C.c_int64]),
    "gsofa_server_release": (C.c_int, [C.c_void_p, C.c_int64]),
    "gsofa_server_retained": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gsofa_server_request_info": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "gsofa_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]),
    "gsofa_server_stop": (C.c_int, [C.c_void_p]),
    "gsofa_digest_fnv1a": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_uint64)]),
    "gsofa_run_noalloc": (C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(Options), C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
}


def lib() -> C.CDLL:
    """The loaded engine library; raises if it was not built."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gsofatpds_amd.build` "
                               "(the engine has no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def _check(rc: int) -> None:
    if rc != OK:
        raise GsofaError(rc, lib().gsofa_last_error().decode(errors="replace"))


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@dataclass
class Pattern:
    """An int32 pattern in the reference's layout: row v is adj[beg[v]:end[v]]."""
    n: int
    beg: np.ndarray
    end: np.ndarray
    adj: np.ndarray

    @classmethod
    def from_rowptr(cls, rowptr: np.ndarray, adj: np.ndarray) -> "Pattern":
        rowptr = np.asarray(rowptr, dtype=np.int64)
        return cls(len(rowptr) - 1, rowptr[:-1].astype(np.int32), rowptr[1:].astype(np.int32),
                   np.ascontiguousarray(adj, dtype=np.int32))

    @classmethod
    def from_dense(cls, dense: np.ndarray) -> "Pattern":
        n = dense.shape[0]
        rows = [np.nonzero(dense[i])[0].astype(np.int32) for i in range(n)]
        ptr = np.zeros(n + 1, np.int64)
        ptr[1:] = np.cumsum([len(r) for r in rows])
        adj = np.concatenate(rows) if n else np.zeros(0, np.int32)
        return cls.from_rowptr(ptr, adj)

    def csr(self) -> CSR:
        self.beg = np.ascontiguousarray(self.beg, dtype=np.int32)
        self.end = np.ascontiguousarray(self.end, dtype=np.int32)
        self.adj = np.ascontiguousarray(self.adj, dtype=np.int32)
        return CSR(self.n, len(self.adj), _i32p(self.beg), _i32p(self.end), _i32p(self.adj))

    @property
    def nnz(self) -> int:
        return int(len(self.adj))


def _take_host_matrix(handle: C.c_void_p) -> Pattern:
    L = lib()
    view = CSR()
    try:
        _check(L.gsofa_host_matrix_csr(handle, C.byref(view)))
        n, nnz = view.n, view.nnz
        beg = np.ctypeslib.as_array(view.beg, (n,)).copy()
        end = np.ctypeslib.as_array(view.end, (n,)).copy()
        adj = np.ctypeslib.as_array(view.adj, (nnz,)).copy() if nnz else np.zeros(0, np.int32)
    finally:
        L.gsofa_host_matrix_free(handle)
    return Pattern(n, beg, end, adj)


def generate_workload(workload: int, member: int = 0) -> Pattern:
    h = C.c_void_p()
    _check(lib().gsofa_generate_workload(workload, member, C.byref(h)))
    return _take_host_matrix(h)


def generate(kind=GEN_STENCIL, dims=2, points=5, nx=0, ny=0, nz=1, n_keep=0, order=ORDER_NATURAL, nd_leaf=64,
             drop_side=DROP_ANY, drop_ppm=0, extra_every=0, avg_degree_x100=600, seed=1) -> Pattern:
    params = (C.c_int64 * 13)(kind, dims, points, nx, ny, nz, n_keep, order, nd_leaf, drop_side, drop_ppm,
                              extra_every, avg_degree_x100)
    h = C.c_void_p()
    _check(lib().gsofa_generate(params, seed, C.byref(h)))
    return _take_host_matrix(h)


def read_ascii(beg_file: str, end_file: str, csr_file: str) -> Pattern:
    h = C.c_void_p()
    _check(lib().gsofa_read_ascii(beg_file.encode(), end_file.encode(), csr_file.encode(), C.byref(h)))
    return _take_host_matrix(h)


def write_ascii(p: Pattern, beg_file: str, end_file: str, csr_file: str) -> None:
    c = p.csr()
    _check(lib().gsofa_write_ascii(C.byref(c), beg_file.encode(), end_file.encode(), csr_file.encode()))


def read_binary(beg_file: str, csr_file: str, index_bytes: int = 8, vert_bytes: int = 8) -> Pattern:
    """The reference's binary CSR pair (row offsets, column ids) at the given integer widths."""
    h = C.c_void_p()
    _check(lib().gsofa_read_binary(beg_file.encode(), csr_file.encode(), index_bytes, vert_bytes, C.byref(h)))
    return _take_host_matrix(h)


def write_binary(p: Pattern, beg_file: str, csr_file: str, index_bytes: int = 8, vert_bytes: int = 8) -> None:
    c = p.csr()
    _check(lib().gsofa_write_binary(C.byref(c), beg_file.encode(), csr_file.encode(), index_bytes, vert_bytes))


def read_matrix_market(path: str) -> Pattern:
    """A square coordinate Matrix Market file's pattern (symmetric kinds expanded)."""
    h = C.c_void_p()
    _check(lib().gsofa_read_matrix_market(path.encode(), C.byref(h)))
    return _take_host_matrix(h)


def read_permutation(path: str, n: int, base: int = 0) -> np.ndarray:
    out = np.zeros(n, np.int32)
    _check(lib().gsofa_read_permutation(path.encode(), n, base, _i32p(out)))
    return out


def permute(p: Pattern, perm_r=None, perm_c=None) -> Pattern:
    """SuperLU_DIST's reordered matrix Pc * Pr * A * Pc^T: B(perm_c[perm_r[i]], perm_c[j]) = A(i, j)."""
    c = p.csr()
    pr = np.ascontiguousarray(perm_r, np.int32) if perm_r is not None else None
    pc = np.ascontiguousarray(perm_c, np.int32) if perm_c is not None else None
    h = C.c_void_p()
    _check(lib().gsofa_permute(C.byref(c), _i32p(pr) if pr is not None else None,
                               _i32p(pc) if pc is not None else None, C.byref(h)))
    return _take_host_matrix(h)


def symbfact_min_id(argv: Sequence[str], rank: int = 0):
    """Mirror of the reference entry: returns (fill_count, time_ms, n_supernodes, edge_checks).

    ``argv`` is the full vector including argv[0] (11 strings), as in the reference.
    """
    L = lib()
    enc = [a.encode() if isinstance(a, str) else a for a in argv]
    arr = (C.c_char_p * len(enc))(*enc)
    fill = C.c_uint64(0)
    t = C.c_double(0.0)
    ns = C.c_int64(0)
    teps = C.c_uint64(0)
    _check(L.gsofa_symbfact_min_id(len(enc), arr, rank, C.byref(fill), C.byref(t), C.byref(ns), C.byref(teps)))
    return fill.value, t.value, ns.value, teps.value


def device_query(device: int = 0) -> dict:
    cus = C.c_int32()
    lb = C.c_int64()
    lc = C.c_int64()
    arch = C.create_string_buffer(128)
    _check(lib().gsofa_device_query(device, C.byref(cus), C.byref(lb), C.byref(lc), arch, 128))
    return {"cus": cus.value, "lds_per_block": lb.value, "lds_per_cu": lc.value, "arch": arch.value.decode()}


def bandwidth_probe(nbytes: int = 1 << 30, iters: int = 10, device: int = 0) -> float:
    """GB/s of a streaming device copy (read + write bytes)."""
    v = C.c_double()
    _check(lib().gsofa_bandwidth_probe(device, nbytes, iters, C.byref(v)))
    return v.value


def device_synchronize(device: int = 0) -> None:
    _check(lib().gsofa_device_synchronize(device))


class Stream:
    """A HIP stream owned by the engine's runtime (use instead of a framework's stream: one
    process must not run two HIP runtimes)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().gsofa_stream_create(device, C.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gsofa_stream_destroy(self.handle)
            self.handle = None

    __del__ = close


def gpu_present() -> bool:
    """True if the engine's HIP runtime sees a device (does not touch any other runtime)."""
    try:
        device_query(0)
        return True
    except (GsofaError, RuntimeError):
        return False


def default_options(**kw) -> Options:
    o = Options()
    lib().gsofa_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, int(v))
    return o


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.handle)  # gsofatpds_amd.Stream


class Plan:
    """One matrix: ``gsofa_plan_*``. Keeps the pattern arrays alive while the plan exists."""

    def __init__(self, pattern: Pattern, **options):
        self.pattern = pattern
        self.options = default_options(**options)
        self._csr = pattern.csr()
        h = C.c_void_p()
        _check(lib().gsofa_plan_create(C.byref(self._csr), C.byref(self.options), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gsofa_plan_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, stream=None) -> "Plan":
        _check(lib().gsofa_run(self._h, _stream_handle(stream)))
        return self

    def sync(self) -> "Plan":
        _check(lib().gsofa_sync(self._h))
        return self

    def result(self) -> Result:
        r = Result()
        _check(lib().gsofa_result_get(self._h, C.byref(r)))
        return r

    def nzu(self) -> np.ndarray:
        out = np.zeros(self.pattern.n, np.int32)
        _check(lib().gsofa_nzu_get(self._h, _i32p(out)))
        return out

    def supernode_flags(self) -> np.ndarray:
        out = np.zeros(self.pattern.n, np.uint8)
        _check(lib().gsofa_supernodes_get(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def stage_timing(self, on: bool) -> "Plan":
        """Record the stage events in later runs (the default) or not (a few microseconds of device
        time per event; stage_times() then raises)."""
        _check(lib().gsofa_plan_stage_timing(self._h, 1 if on else 0))
        return self

    def stage_times(self) -> dict:
        ms = (C.c_double * 4)()
        _check(lib().gsofa_stage_times(self._h, ms))
        return {"prep_ms": ms[0], "light_ms": ms[1], "heavy_ms": ms[2], "supernodes_ms": ms[3]}

    def work_split(self) -> dict:
        w, r, e = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().gsofa_work_split(self._h, C.byref(w), C.byref(r), C.byref(e)))
        return {"heavy_edge_checks": w.value, "heavy_reach": r.value, "heavy_entries": e.value}

    def config(self) -> dict:
        """Launch configuration in use (after any autotuning)."""
        w = (C.c_int32 * 4)()
        _check(lib().gsofa_plan_config(self._h, w))
        return {"heavy_threads": w[0], "heavy_wgs_per_cu": w[1], "chain_block": w[2], "queue_capacity": w[3]}

    def executed_work(self) -> dict:
        w = (C.c_uint64 * 6)()
        _check(lib().gsofa_executed_work(self._h, w))
        return {"heavy_edges": w[0], "heavy_rows": w[1], "heavy_out": w[2],
                "light_edges": w[3], "light_rows": w[4], "light_out": w[5]}

    def early_top(self) -> dict:
        """The last run's early launch of the top separator's items: item count (0: none), device
        time (0 unless the run recorded stage events) and executed work."""
        items, ms, w = C.c_int64(), C.c_double(), (C.c_uint64 * 3)()
        _check(lib().gsofa_early_top(self._h, C.byref(items), C.byref(ms), w))
        return {"items": items.value, "ms": ms.value, "edges": w[0], "rows": w[1], "out": w[2]}

    def structure(self):
        """(u_rowptr, u_col, l_colptr, l_row): strict U by rows, strict L by columns, sorted."""
        nu, nl = C.c_int64(), C.c_int64()
        _check(lib().gsofa_structure_size(self._h, C.byref(nu), C.byref(nl)))
        n = self.pattern.n
        up = np.zeros(n + 1, np.int64)
        lp = np.zeros(n + 1, np.int64)
        ui = np.zeros(max(nu.value, 1), np.int32)
        li = np.zeros(max(nl.value, 1), np.int32)
        _check(lib().gsofa_structure_get(self._h, up.ctypes.data_as(C.POINTER(C.c_int64)), _i32p(ui),
                                         lp.ctypes.data_as(C.POINTER(C.c_int64)), _i32p(li)))
        return up, ui[:nu.value], lp, li[:nl.value]

    def supernode_partition(self):
        """(xsup, supno): supernode k is rows [xsup[k], xsup[k+1]); supno[i] is row i's supernode."""
        ns = C.c_int64()
        _check(lib().gsofa_supernode_partition(self._h, C.byref(ns), None, None))
        xsup = np.zeros(ns.value + 1, np.int32)
        supno = np.zeros(self.pattern.n, np.int32)
        _check(lib().gsofa_supernode_partition(self._h, C.byref(ns), _i32p(xsup), _i32p(supno)))
        return xsup, supno

    def supernodal_structure(self):
        """(xsup, xlsub, lsub, xusub, usub): per-supernode L panel rows and U panel columns."""
        ns, nl, nu = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().gsofa_supernodal_structure_size(self._h, C.byref(ns), C.byref(nl), C.byref(nu)))
        xsup = np.zeros(ns.value + 1, np.int32)
        xl = np.zeros(ns.value + 1, np.int64)
        xu = np.zeros(ns.value + 1, np.int64)
        ls = np.zeros(max(nl.value, 1), np.int32)
        us = np.zeros(max(nu.value, 1), np.int32)
        _check(lib().gsofa_supernodal_structure_get(self._h, _i32p(xsup), xl.ctypes.data_as(C.POINTER(C.c_int64)),
                                                    _i32p(ls), xu.ctypes.data_as(C.POINTER(C.c_int64)), _i32p(us)))
        return xsup, xl, ls[:nl.value], xu, us[:nu.value]


def supernodal_from_structure(flags, u_rowptr, u_col, l_colptr, l_row):
    """``gsofa_supernodal_from_structure`` on host arrays (no device): (xsup, xlsub, lsub, xusub, usub)."""
    n = len(flags)
    fl = np.ascontiguousarray(flags, np.uint8)
    up = np.ascontiguousarray(u_rowptr, np.int64)
    lp = np.ascontiguousarray(l_colptr, np.int64)
    ui = np.ascontiguousarray(u_col, np.int32) if len(u_col) else np.zeros(1, np.int32)
    li = np.ascontiguousarray(l_row, np.int32) if len(l_row) else np.zeros(1, np.int32)
    p64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    u8 = fl.ctypes.data_as(C.POINTER(C.c_uint8))
    ns, nl, nu = C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib().gsofa_supernodal_from_structure(n, u8, p64(up), _i32p(ui), p64(lp), _i32p(li), C.byref(ns),
                                                 C.byref(nl), C.byref(nu), None, None, None, None, None))
    xsup = np.zeros(ns.value + 1, np.int32)
    xl = np.zeros(ns.value + 1, np.int64)
    xu = np.zeros(ns.value + 1, np.int64)
    ls = np.zeros(max(nl.value, 1), np.int32)
    us = np.zeros(max(nu.value, 1), np.int32)
    _check(lib().gsofa_supernodal_from_structure(n, u8, p64(up), _i32p(ui), p64(lp), _i32p(li), C.byref(ns),
                                                 C.byref(nl), C.byref(nu), _i32p(xsup), p64(xl), _i32p(ls), p64(xu),
                                                 _i32p(us)))
    return xsup, xl, ls[:nl.value], xu, us[:nu.value]


NOALLOC_TOTALS = ("fill_count", "nnz_F", "nnz_L", "nnz_U", "edge_checks", "reach", "supernodes",
                  "supernodes_counted", "input_error")


def run_noalloc(n: int, nnz: int, rowptr: int, adj: int, nzu: int, sup_flags: int, totals: int, stream=None,
                **options) -> None:
    """``gsofa_run_noalloc``: enqueue the counts-only run of one pattern without allocating device memory.

    Every pointer is a device address: ``rowptr`` (n + 1 int32) and ``adj`` (nnz int32) hold rows of
    strictly increasing column ids; ``nzu`` (n int32), ``sup_flags`` (n uint8) and ``totals``
    (``len(NOALLOC_TOTALS)`` uint64, in that order) receive the results on ``stream``. A pattern that
    breaks the row rule leaves ``input_error`` set and the counts of an empty pattern."""
    o = default_options(**options)
    _check(lib().gsofa_run_noalloc(int(n), int(nnz), rowptr, adj, C.byref(o), nzu, sup_flags, totals,
                                   _stream_handle(stream)))


class Batch:
    """Many matrices in one captured graph: ``gsofa_batch_*``."""

    def __init__(self, patterns: Sequence[Pattern], **options):
        self.patterns = list(patterns)
        self.options = default_options(**options)
        self._csrs = (CSR * len(self.patterns))(*[p.csr() for p in self.patterns])
        h = C.c_void_p()
        _check(lib().gsofa_batch_create(self._csrs, len(self.patterns), C.byref(self.options), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gsofa_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def config(self) -> dict:
        """Launch configuration in use (after any autotuning)."""
        w = (C.c_int32 * 4)()
        _check(lib().gsofa_batch_config(self._h, w))
        return {"heavy_threads": w[0], "heavy_wgs_per_cu": w[1], "chain_block": w[2], "queue_capacity": w[3]}

    def run(self, stream=None) -> "Batch":
        _check(lib().gsofa_batch_run(self._h, _stream_handle(stream)))
        return self

    def sync(self) -> "Batch":
        _check(lib().gsofa_batch_sync(self._h))
        return self

    def result(self, i: int) -> Result:
        r = Result()
        _check(lib().gsofa_batch_result_get(self._h, i, C.byref(r)))
        return r

    def nzu(self, i: int) -> np.ndarray:
        out = np.zeros(self.patterns[i].n, np.int32)
        _check(lib().gsofa_batch_nzu_get(self._h, i, _i32p(out)))
        return out

    def supernode_partition(self, i: int):
        """(xsup, supno) of member i, as Plan.supernode_partition."""
        ns = C.c_int64()
        _check(lib().gsofa_batch_supernode_partition(self._h, i, C.byref(ns), None, None))
        xsup = np.zeros(ns.value + 1, np.int32)
        supno = np.zeros(self.patterns[i].n, np.int32)
        _check(lib().gsofa_batch_supernode_partition(self._h, i, C.byref(ns), _i32p(xsup), _i32p(supno)))
        return xsup, supno

    def supernode_flags(self, i: int) -> np.ndarray:
        out = np.zeros(self.patterns[i].n, np.uint8)
        _check(lib().gsofa_batch_supernodes_get(self._h, i, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out


class ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("cu_share", C.c_int32), ("max_group", C.c_int32)]


class Server:
    """Continuous-batching request loop: ``gsofa_server_*``."""

    def __init__(self, slots: int = 4, hbm_budget_bytes: int = 0, device: int = 0, cu_share: int = 0,
                 max_group: int = 0):
        o = ServerOptions()
        lib().gsofa_server_options_default(C.byref(o))
        o.slots, o.hbm_budget_bytes, o.device, o.cu_share = slots, hbm_budget_bytes, device, cu_share
        if max_group:
            o.max_group = max_group
        h = C.c_void_p()
        _check(lib().gsofa_server_start(C.byref(o), C.byref(h)))
        self._h = h

    def submit(self, pattern: Pattern, **options) -> int:
        c = pattern.csr()
        o = default_options(**options)
        rid = C.c_int64()
        _check(lib().gsofa_server_submit(self._h, C.byref(c), C.byref(o), C.byref(rid)))
        return rid.value

    def poll(self, rid: int) -> Optional[Result]:
        r = Result()
        rc = lib().gsofa_server_poll(self._h, rid, C.byref(r))
        if rc == ERR_NOT_READY:
            return None
        _check(rc)
        return r

    def wait(self, rid: int, timeout_s: float = -1.0) -> Result:
        r = Result()
        _check(lib().gsofa_server_wait(self._h, rid, timeout_s, C.byref(r)))
        return r

    def cancel(self, rid: int) -> None:
        _check(lib().gsofa_server_cancel(self._h, rid))

    def info(self, rid: int) -> dict:
        lat, q = C.c_double(), C.c_double()
        dz, ds = C.c_uint64(), C.c_uint64()
        _check(lib().gsofa_server_request_info(self._h, rid, C.byref(lat), C.byref(q), C.byref(dz), C.byref(ds)))
        return {"latency_ms": lat.value, "queue_ms": q.value, "nzu_digest": dz.value, "supernode_digest": ds.value}

    def release(self, rid: int) -> None:
        _check(lib().gsofa_server_release(self._h, rid))

    def retained(self) -> dict:
        n, b = C.c_int64(), C.c_int64()
        _check(lib().gsofa_server_retained(self._h, C.byref(n), C.byref(b)))
        return {"requests": n.value, "host_bytes": b.value}

    def stats(self) -> dict:
        c, pn, iu, b = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        r, pr = C.c_int32(), C.c_int32()
        _check(lib().gsofa_server_stats(self._h, C.byref(c), C.byref(r), C.byref(pr), C.byref(pn), C.byref(iu),
                                        C.byref(b)))
        return {"completed": c.value, "running": r.value, "peak_running": pr.value, "pending": pn.value,
                "hbm_in_use": iu.value, "hbm_budget": b.value}

    def stop(self) -> None:
        if getattr(self, "_h", None):
            lib().gsofa_server_stop(self._h)
            self._h = None

    __del__ = stop


def fnv1a(data: bytes) -> int:
    """The digest the server reports for nzU and the supernode flags (gsofa_digest_fnv1a)."""
    out = C.c_uint64()
    _check(lib().gsofa_digest_fnv1a(data, len(data), C.byref(out)))
    return out.value


def fnv1a_array(a: np.ndarray) -> int:
    """fnv1a of an array's bytes (int32 nzU, uint8 flags), without a copy."""
    a = np.ascontiguousarray(a)
    out = C.c_uint64()
    _check(lib().gsofa_digest_fnv1a(a.ctypes.data_as(C.c_void_p), a.nbytes, C.byref(out)))
    return out.value
