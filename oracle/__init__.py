"""ORACLE - test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's CPU baseline).

Python bindings of ``oracle/oracle.cpp`` (the host C++ restatement of the reference's
traversal and supernode rule) plus two independent checkers written from the mathematics,
used to pin the restatement:

  * ``dense_elimination``: structural Gaussian elimination without pivoting on a dense boolean
    matrix (the fill-path theorem's definition of F, no cancellation);
  * ``etree_cholesky_fill``: symbolic Cholesky of a symmetric pattern through the elimination
    tree (row subtrees), for the symmetric cross-check fill(L+U) = 2 fill(L).

The engine never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OracleResult(C.Structure):
    _fields_ = [("fill_count", C.c_uint64), ("nnz_F", C.c_uint64), ("W", C.c_uint64), ("R", C.c_uint64),
                ("expansions", C.c_uint64), ("edge_checks", C.c_uint64), ("nsup", C.c_int64)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        h.oracle_symbfact.restype = C.c_int
        h.oracle_sources.restype = C.c_int
        h.oracle_max_threads.restype = C.c_int
        h.oracle_sources_w.restype = C.c_int
        _lib = h
    return _lib


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _packed(pattern):
    """Packed CSR (rowptr int64, adj int32) of a pattern with beg/end arrays."""
    beg = np.asarray(pattern.beg, np.int64)
    end = np.asarray(pattern.end, np.int64)
    adj = np.asarray(pattern.adj, np.int32)
    n = len(beg)
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum(end - beg)
    idx = np.concatenate([np.arange(b, e) for b, e in zip(beg, end)]) if n else np.zeros(0, np.int64)
    return ptr, adj[idx] if len(idx) else np.zeros(0, np.int32)


def symbfact(pattern, chunk=128, threads=0, structure=False, stats=True):
    """Runs the restatement. Returns a dict with fill, nnz_F, W, R, nsup (counted), nzu,
    sup_flag and, with structure=True, the F rows (rowptr, cols; diagonal included)."""
    L = lib()
    n = int(pattern.n)
    beg = np.ascontiguousarray(pattern.beg, np.int32)
    end = np.ascontiguousarray(pattern.end, np.int32)
    adj = np.ascontiguousarray(pattern.adj, np.int32)
    if len(adj) == 0:
        adj = np.zeros(1, np.int32)
    res = OracleResult()
    nzu = np.zeros(n, np.int32)
    flags = np.zeros(n, np.uint8)
    row_nnz = np.zeros(n, np.int64)
    rc = L.oracle_symbfact(C.c_int32(n), _p32(beg), _p32(end), _p32(adj), C.c_int32(chunk), C.c_int32(threads),
                           C.c_int32(1 if stats else 0), _p32(nzu), row_nnz.ctypes.data_as(C.POINTER(C.c_int64)),
                           None, None, flags.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(res))
    if rc != 0:
        raise RuntimeError("oracle_symbfact failed")
    out = dict(fill=res.fill_count, nnz_F=res.nnz_F, W=res.W, R=res.R, nsup=res.nsup, expansions=res.expansions,
               edge_checks=res.edge_checks, nzu=nzu, sup_flag=flags)
    if structure:
        ptr = np.zeros(n + 1, np.int64)
        ptr[1:] = np.cumsum(row_nnz)
        cols = np.zeros(max(int(ptr[-1]), 1), np.int32)
        res2 = OracleResult()
        rc = L.oracle_symbfact(C.c_int32(n), _p32(beg), _p32(end), _p32(adj), C.c_int32(chunk), C.c_int32(threads),
                               C.c_int32(0), None, None, ptr.ctypes.data_as(C.POINTER(C.c_int64)), _p32(cols), None,
                               C.byref(res2))
        if rc != 0:
            raise RuntimeError("oracle_symbfact failed")
        out["F_rowptr"] = ptr
        out["F_cols"] = cols[: int(ptr[-1])]
    return out


def sources(pattern, srcs, threads=0):
    """Traversal of a subset of sources: (fills, edge checks). Used for bounded CPU samples."""
    L = lib()
    beg = np.ascontiguousarray(pattern.beg, np.int32)
    end = np.ascontiguousarray(pattern.end, np.int32)
    adj = np.ascontiguousarray(pattern.adj, np.int32)
    s = np.ascontiguousarray(srcs, np.int32)
    f = C.c_uint64()
    c = C.c_uint64()
    rc = L.oracle_sources(C.c_int32(int(pattern.n)), _p32(beg), _p32(end), _p32(adj), _p32(s), C.c_int64(len(s)),
                          C.c_int32(threads), C.byref(f), C.byref(c))
    if rc != 0:
        raise RuntimeError("oracle_sources failed")
    return f.value, c.value


def sources_work(pattern, srcs, threads=0):
    """Schedule-independent (W, |R|) summed over a subset of sources."""
    L = lib()
    beg = np.ascontiguousarray(pattern.beg, np.int32)
    end = np.ascontiguousarray(pattern.end, np.int32)
    adj = np.ascontiguousarray(pattern.adj, np.int32)
    s = np.ascontiguousarray(srcs, np.int32)
    w = C.c_uint64()
    r = C.c_uint64()
    rc = L.oracle_sources_w(C.c_int32(int(pattern.n)), _p32(beg), _p32(end), _p32(adj), _p32(s), C.c_int64(len(s)),
                            C.c_int32(threads), C.byref(w), C.byref(r))
    if rc != 0:
        raise RuntimeError("oracle_sources_w failed")
    return w.value, r.value


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def split_structure(n, F_rowptr, F_cols):
    """F rows (diagonal included) -> strict U by rows and strict L by columns, sorted."""
    u_rows, l_cols = [], [[] for _ in range(n)]
    for i in range(n):
        row = F_cols[F_rowptr[i]:F_rowptr[i + 1]]
        u_rows.append(row[row > i])
        for m in row[row < i]:
            l_cols[m].append(i)
    u_ptr = np.zeros(n + 1, np.int64)
    u_ptr[1:] = np.cumsum([len(r) for r in u_rows])
    l_ptr = np.zeros(n + 1, np.int64)
    l_ptr[1:] = np.cumsum([len(c) for c in l_cols])
    u_idx = np.concatenate(u_rows).astype(np.int32) if u_ptr[-1] else np.zeros(0, np.int32)
    l_idx = np.array([x for c in l_cols for x in sorted(c)], np.int32)
    return u_ptr, u_idx, l_ptr, l_idx


def split_structure_fast(n, F_rowptr, F_cols):
    """split_structure for large F (vectorised): strict U by rows, strict L by columns, sorted."""
    ptr = np.asarray(F_rowptr, np.int64)
    cols = np.asarray(F_cols, np.int32)
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(ptr))
    up = cols > rows
    u_idx = cols[up]
    u_ptr = np.zeros(n + 1, np.int64)
    u_ptr[1:] = np.cumsum(np.bincount(rows[up], minlength=n))
    lo = cols < rows
    lc, lr = cols[lo], rows[lo]
    del rows, up, lo
    order = np.lexsort((lr, lc))  # by column, then row
    l_idx = lr[order]
    l_ptr = np.zeros(n + 1, np.int64)
    l_ptr[1:] = np.cumsum(np.bincount(lc, minlength=n))
    return u_ptr, u_idx, l_ptr, l_idx


def structure_digests(u_ptr, u_idx, l_ptr, l_idx):
    """SHA-256 of strict U by rows (int64 offsets, then int32 columns) and of strict L by columns
    (int64 offsets, then int32 rows): the form gsofa_structure_get returns."""
    import hashlib
    hu = hashlib.sha256()
    hu.update(np.ascontiguousarray(u_ptr, np.int64).tobytes())
    hu.update(np.ascontiguousarray(u_idx, np.int32).tobytes())
    hl = hashlib.sha256()
    hl.update(np.ascontiguousarray(l_ptr, np.int64).tobytes())
    hl.update(np.ascontiguousarray(l_idx, np.int32).tobytes())
    return hu.hexdigest(), hl.hexdigest()


# ---- independent checkers (from the mathematics, not from the reference) --------------------

def dense_elimination(dense: np.ndarray) -> np.ndarray:
    """Structure of L+U from Gaussian elimination without pivoting, no numerical cancellation:
    for k in order, every row i > k with F[i,k] receives row k's entries right of k."""
    F = dense.astype(bool).copy()
    n = F.shape[0]
    for k in range(n):
        rows = np.nonzero(F[k + 1:, k])[0] + k + 1
        if len(rows):
            F[np.ix_(rows, np.arange(k + 1, n))] |= F[k, k + 1:]
    return F


def etree_cholesky_fill(n, rowptr, adj):
    """Strict-lower nonzeros of the Cholesky factor of a symmetric pattern (Liu's row-subtree
    algorithm: the elimination tree by path compression, then each row's reach in the tree)."""
    parent = np.full(n, -1, np.int64)
    ancestor = np.full(n, -1, np.int64)
    for i in range(n):
        for j in adj[rowptr[i]:rowptr[i + 1]]:
            j = int(j)
            while j != -1 and j < i:
                nxt = ancestor[j]
                ancestor[j] = i
                if nxt == -1:
                    parent[j] = i
                j = nxt
    count = 0
    mark = np.full(n, -1, np.int64)
    for i in range(n):
        mark[i] = i
        for j in adj[rowptr[i]:rowptr[i + 1]]:
            j = int(j)
            while j < i and mark[j] != i:
                mark[j] = i
                count += 1
                j = int(parent[j])
                if j == -1:
                    break
    return count
