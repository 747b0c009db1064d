"""Pins of the oracle (CPU only): the reference's known answer, brute force, closed forms,
the symmetric elimination-tree cross-check, hand-worked supernode cases and invariances."""
import json
import os

import numpy as np
import pytest

import oracle
import gsofatpds_amd as g
from conftest import GOLDEN


def load_ds():
    z = np.load(os.path.join(GOLDEN, "ds_11948.npz"))
    return g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])


def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


def dense_to_pattern(D):
    return g.Pattern.from_dense(D)


def oracle_dense_F(p):
    o = oracle.symbfact(p, structure=True, stats=False)
    n = p.n
    F = np.zeros((n, n), bool)
    for i in range(n):
        F[i, o["F_cols"][o["F_rowptr"][i]:o["F_rowptr"][i + 1]]] = True
    return F, o


# ---- the reference's known answer ------------------------------------------------------------

def test_reference_dataset_fill_count():
    """SpaceOptimization/with_supernodes/space_optimization_log_1_264_GPU_warp_centric.dat:2 records
    980136 fill-ins for the shipped 11,948-row input."""
    ds = load_ds()
    assert ds.n == 11948 and ds.nnz == 149090
    o = oracle.symbfact(ds, chunk=128)
    assert o["fill"] == 980136
    assert o["nnz_F"] == 980136 + 149090  # every row of the input holds its diagonal


def test_reference_dataset_matches_committed_golden():
    gd = golden()["ds_11948"]["oracle_S128"]
    o = oracle.symbfact(load_ds(), chunk=128)
    for k in ("fill", "nnz_F", "W", "R", "nsup"):
        assert o[k] == gd[k], k


def test_reference_dataset_symmetric_etree():
    """The dataset is structurally symmetric: fill(L+U) = 2 fill(L) from symbolic Cholesky."""
    ds = load_ds()
    ptr, adj = oracle._packed(ds)
    n = ds.n
    lower_a = sum(int(np.sum(adj[ptr[i]:ptr[i + 1]] < i)) for i in range(n))
    count = oracle.etree_cholesky_fill(n, ptr, adj)
    assert 2 * (count - lower_a) == 980136


# ---- brute force: dense structural elimination ------------------------------------------------

@pytest.mark.parametrize("n,density,seed", [(8, 0.3, 1), (40, 0.08, 2), (64, 0.05, 3), (90, 0.03, 4),
                                            (120, 0.02, 5), (120, 0.06, 6), (200, 0.015, 7)])
def test_oracle_equals_dense_elimination(n, density, seed):
    rng = np.random.default_rng(seed)
    D = rng.random((n, n)) < density
    np.fill_diagonal(D, True)
    p = dense_to_pattern(D)
    F, o = oracle_dense_F(p)
    E = dense_elimination = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)
    assert np.array_equal(F, E)
    off = ~np.eye(n, dtype=bool)
    assert o["fill"] == int(np.sum(E & ~D & off))
    # nzU[i] counts the upper part of F row i
    assert np.array_equal(o["nzu"], np.triu(E, 1).sum(axis=1))


def test_oracle_missing_diagonal_and_empty_rows():
    rng = np.random.default_rng(11)
    D = rng.random((70, 70)) < 0.05
    D[5, :] = False  # an empty row
    D[:, 9] = False
    p = dense_to_pattern(D)
    F, o = oracle_dense_F(p)
    E = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)  # the diagonal always belongs to F
    assert np.array_equal(F, E)


# ---- closed forms ---------------------------------------------------------------------------

def tridiagonal(n):
    D = np.eye(n, dtype=bool) | np.eye(n, k=1, dtype=bool) | np.eye(n, k=-1, dtype=bool)
    return D


def test_tridiagonal_no_fill_and_exact_work():
    n = 300
    p = dense_to_pattern(tridiagonal(n))
    o = oracle.symbfact(p)
    assert o["fill"] == 0
    # R_i = {0..i-1}: the chain below i is reachable through vertices < i
    assert o["R"] == n * (n - 1) // 2
    deg = np.full(n, 3)
    deg[0] = deg[-1] = 2
    assert o["W"] == int(sum(deg[i] + deg[:i].sum() for i in range(n)))


def test_band_no_fill():
    n, k = 200, 4
    D = np.zeros((n, n), bool)
    for d in range(-k, k + 1):
        D |= np.eye(n, k=d, dtype=bool)
    assert oracle.symbfact(dense_to_pattern(D))["fill"] == 0


def test_arrow_dense_first_row_fills_completely():
    n = 100
    D = np.eye(n, dtype=bool)
    D[0, :] = D[:, 0] = True
    o = oracle.symbfact(dense_to_pattern(D))
    assert o["fill"] == n * n - (3 * n - 2)
    assert o["nnz_F"] == n * n


def test_arrow_dense_last_row_no_fill():
    n = 100
    D = np.eye(n, dtype=bool)
    D[-1, :] = D[:, -1] = True
    assert oracle.symbfact(dense_to_pattern(D))["fill"] == 0


def test_grid_natural_order_fills_the_envelope():
    """k x k 5-point grid in natural order: L+U is exactly the envelope, (i, j) in F iff
    min(i, j) >= f[max(i, j)] with f[v] the first column of row v (a path along the grid row
    below reaches every vertex between f[v] and v through smaller vertices)."""
    k = 12
    n = k * k
    D = np.eye(n, dtype=bool)
    for v in range(n):
        x, y = v % k, v // k
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            if 0 <= x + dx < k and 0 <= y + dy < k:
                D[v, (y + dy) * k + x + dx] = True
    F, o = oracle_dense_F(dense_to_pattern(D))
    first = D.argmax(axis=1)
    i, j = np.indices((n, n))
    assert np.array_equal(F, np.minimum(i, j) >= first[np.maximum(i, j)])
    assert o["fill"] > 0


# ---- symmetric cross-check -------------------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_symmetric_random_equals_twice_cholesky_fill(seed):
    rng = np.random.default_rng(seed)
    n = 400
    D = rng.random((n, n)) < 0.006
    D = D | D.T
    np.fill_diagonal(D, True)
    p = dense_to_pattern(D)
    o = oracle.symbfact(p)
    ptr, adj = oracle._packed(p)
    lower_a = int(np.tril(D, -1).sum())
    assert o["fill"] == 2 * (oracle.etree_cholesky_fill(n, ptr, adj) - lower_a)


# ---- supernode rule: hand-worked cases ----------------------------------------------------------
# Rule (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:1790-2123): start if
# (r mod S) mod 128 == 0 or nzU[r-1] - nzU[r] != 1; a pass-through row extends iff (r, start) in F.

def test_supernodes_dense_matrix():
    n = 300
    p = dense_to_pattern(np.ones((n, n), bool))
    for S, expect in ((128, 3), (256, 3), (64, 5), (1024, 3)):
        # only the forced starts count: rows with (r mod S) mod 128 == 0
        o = oracle.symbfact(p, chunk=S)
        assert o["nsup"] == expect, S
        assert o["nsup"] == sum(1 for r in range(n) if (r % S) % 128 == 0)


def test_supernodes_diagonal_every_row_starts():
    n = 333
    o = oracle.symbfact(dense_to_pattern(np.eye(n, dtype=bool)), chunk=128)
    assert o["nsup"] == n  # nzU = 0 everywhere: the difference is never 1


def test_supernodes_upper_triangle_every_pass_through_fails():
    n = 260
    D = np.triu(np.ones((n, n), bool))
    o = oracle.symbfact(dense_to_pattern(D), chunk=128)
    # nzU[r] = n-1-r so every non-forced row passes the first filter, but F has no strict
    # lower entries, so each one fails the second filter and counts a supernode
    assert o["fill"] == 0
    assert o["nsup"] == n
    assert int(o["sup_flag"].sum()) == n


def test_supernodes_lower_bidiagonal_plus_upper_chain():
    """A(i,i+1) and A(i+1,i): tridiagonal. nzU = 1 except the last row (0): only the last row
    passes the first filter; (n-1, n-2) is in A, so it extends the run of row n-2."""
    n = 100
    o = oracle.symbfact(dense_to_pattern(tridiagonal(n)), chunk=128)
    assert o["nsup"] == n - 1


# ---- invariances -----------------------------------------------------------------------------

def test_invariance_chunk_and_threads():
    p = g.generate_workload(g.WL_GRID2D_4096)
    base = oracle.symbfact(p, chunk=128, threads=1)
    for S in (32, 64, 256, 1024):
        for t in (1, 3, 8):
            o = oracle.symbfact(p, chunk=S, threads=t)
            for k in ("fill", "nnz_F", "W", "R"):
                assert o[k] == base[k]
            assert np.array_equal(o["nzu"], base["nzu"])
    a = oracle.symbfact(p, chunk=256, threads=1)
    b = oracle.symbfact(p, chunk=256, threads=8)
    assert np.array_equal(a["sup_flag"], b["sup_flag"])


def test_grid_workload_matches_committed_golden():
    import hashlib
    gd = golden()["grid2d_4096"]
    p = g.generate_workload(g.WL_GRID2D_4096)
    o = oracle.symbfact(p, chunk=128, structure=True)
    for k in ("fill", "nnz_F", "W", "R", "nsup"):
        assert o[k] == gd["oracle_S128"][k]
    h = hashlib.sha256()
    h.update(o["F_rowptr"].astype(np.int64).tobytes())
    h.update(o["F_cols"].astype(np.int32).tobytes())
    assert h.hexdigest() == gd["F_sha256"]
    assert hashlib.sha256(o["nzu"].astype(np.int32).tobytes()).hexdigest() == gd["nzu_sha256"]
    assert hashlib.sha256(o["sup_flag"].astype(np.uint8).tobytes()).hexdigest() == gd["sup_flag_sha256"]


def test_planted_mistake_is_caught():
    """A plausible mistake (a dropped fill test term: marking every reached vertex as fill)
    planted on the CPU must disagree with dense elimination."""
    rng = np.random.default_rng(5)
    n = 80
    D = rng.random((n, n)) < 0.05
    np.fill_diagonal(D, True)
    E = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)
    # wrong rule: every vertex reached through vertices < i is in row i (no m > max-intermediate test)
    wrong = np.zeros_like(E)
    for i in range(n):
        seen = {i}
        stack = [m for m in np.nonzero(D[i])[0] if m < i]
        reach = set(stack)
        while stack:
            v = stack.pop()
            for m in np.nonzero(D[v])[0]:
                if m < i and m not in reach:
                    reach.add(m)
                    stack.append(m)
        wrong[i, list(reach | seen)] = True
        for v in reach | {i}:
            wrong[i, np.nonzero(D[v])[0]] |= np.nonzero(D[v])[0] > i
    assert not np.array_equal(wrong, E)
