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


# ---- supernode rule: the walk after a failed pass-through ----------------------------------------
# In phase 2 (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:2028-2049) a walk from
# start s tests every following pass-through row r against the same s: (r, s) in F extends the
# run (:2034-2038), otherwise one more supernode is counted (:2047) and the walk goes on, still
# comparing against s (no start or end is recorded, :2048-2049). The cases below separate that
# from the plausible misreading "the failed row starts a new run".

def rows_to_dense(n, rows):
    D = np.zeros((n, n), bool)
    for i, r in enumerate(rows):
        D[i, r] = True
    return D


def test_supernode_failed_pass_through_keeps_the_start():
    """n = 3, rows {0,1,2}, {1,2}, {1,2}. nzU = 2, 1, 0, so rows 1 and 2 pass the first filter.
    (1,0) is not in F (row 1 has no lower entry) and (2,0) is not in F (row 2's only lower entry
    is 1, and no vertex below 0 exists), so both fail against start 0: 3 supernodes.
    Restarting at row 1 would extend row 2, since (2,1) is in A: 2 supernodes."""
    o = oracle.symbfact(dense_to_pattern(rows_to_dense(3, [[0, 1, 2], [1, 2], [1, 2]])), chunk=128)
    assert list(o["nzu"]) == [2, 1, 0]
    assert o["nsup"] == 3
    assert list(o["sup_flag"]) == [1, 1, 1]


def test_supernode_row_after_a_failure_extends_against_the_start():
    """n = 3, rows {0,1,2}, {1,2}, {0,2}: row 1 fails against start 0, row 2 holds (2,0) in A and
    extends the run of row 0: 2 supernodes, flags 1, 1, 0."""
    o = oracle.symbfact(dense_to_pattern(rows_to_dense(3, [[0, 1, 2], [1, 2], [0, 2]])), chunk=128)
    assert o["nsup"] == 2
    assert list(o["sup_flag"]) == [1, 1, 0]


def test_supernode_walk_dense_upper_hand_worked():
    """n = 6, the whole upper triangle present, lower entries: row 2 {1}, row 3 {0}, row 5 {2}.
    nzU[r] = 5 - r, so every row after 0 is a pass-through. With U dense, (r, s) in F iff row r
    has a lower entry x <= s (r -> x -> s, x < s, or x = s). Against start 0: rows 1, 2, 4 and 5
    fail, row 3 extends: 5 supernodes, flags 1 1 1 0 1 1. (Restarting at each failed row would give
    3: rows 2, 3 and 5 would extend runs of rows 1, 1 and 4.)"""
    n = 6
    D = np.triu(np.ones((n, n), bool))
    D[2, 1] = D[3, 0] = D[5, 2] = True
    o = oracle.symbfact(dense_to_pattern(D), chunk=128)
    assert list(o["nzu"]) == [5, 4, 3, 2, 1, 0]
    assert o["nsup"] == 5
    assert list(o["sup_flag"]) == [1, 1, 1, 0, 1, 1]


def two_phase_supernodes(F, nzu, S, restart_on_failure=False):
    """Supernode count and increment flags, written from the reference's two phases:
    phase 1 (...cuh:1833-1950): the 128 threads of one block take the chunk's offsets t, t+128, ...;
    thread 0's rows are starts (:1879), another row is a start when nzU[r-1] - nzU[r] != 1
    (:1928), else it gets a pass-through flag (:1943). Phase 2 (:1961-2111): from each start, walk
    r = s+1, s+2, ... while r carries a flag (flags of the next chunk are not set yet, so the walk
    ends at the chunk's end): (r, s) in F extends the run, else one more supernode.
    restart_on_failure plants the misreading: a failed row becomes the start of the walk."""
    n = len(nzu)
    flags = np.zeros(n, np.uint8)
    count = 0
    for c0 in range(0, n, S):
        c1 = min(c0 + S, n)
        starts, passing = [], np.zeros(c1 - c0, bool)
        for off in range(c1 - c0):
            r = c0 + off
            if off % 128 == 0 or nzu[r - 1] - nzu[r] != 1:
                starts.append(r)
            else:
                passing[off] = True
        for s in starts:
            count += 1
            flags[s] = 1
            base = s
            r = s + 1
            while r < c1 and passing[r - c0]:
                if not F[r, base]:
                    count += 1
                    flags[r] = 1
                    if restart_on_failure:
                        base = r
                r += 1
    return count, flags


def pass_through_family(n, p_lower, seed):
    """Whole upper triangle plus random lower entries: every row passes the first filter, and
    (r, s) in F depends on row r's lowest entry, so both outcomes of the second filter occur."""
    rng = np.random.default_rng(seed)
    D = np.triu(np.ones((n, n), bool)) | np.tril(rng.random((n, n)) < p_lower, -1)
    return D


@pytest.mark.parametrize("S", [64, 128, 256])
@pytest.mark.parametrize("family,n,param,seed", [("upper", 256, 0.01, 1), ("upper", 200, 0.03, 2),
                                                  ("upper", 131, 0.005, 3), ("random", 256, 0.02, 4),
                                                  ("random", 180, 0.05, 5), ("band", 256, 6, 6)])
def test_supernodes_equal_two_phase_brute_force(S, family, n, param, seed):
    """The oracle's supernodes against the two-phase restatement run on F from dense elimination
    and nzU counted from that F. The inputs exercise failed pass-throughs, extensions after a
    failure and, for S = 256, the forced start at offset 128; the planted restart variant must
    disagree with the oracle wherever a failure is followed by a row of the same walk."""
    rng = np.random.default_rng(seed)
    if family == "upper":
        D = pass_through_family(n, param, seed)
    elif family == "random":
        D = rng.random((n, n)) < param
        np.fill_diagonal(D, True)
    else:  # an upper band of random width per row and a sparse lower part
        D = np.eye(n, dtype=bool) | np.tril(rng.random((n, n)) < 0.01, -1)
        for i in range(n):
            D[i, i + 1:i + 1 + int(rng.integers(0, param))] = True
    E = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)
    nzu = np.triu(E, 1).sum(axis=1)
    count, flags = two_phase_supernodes(E, nzu, S)
    o = oracle.symbfact(dense_to_pattern(D), chunk=S)
    assert np.array_equal(o["nzu"], nzu)
    assert o["nsup"] == count
    assert np.array_equal(o["sup_flag"], flags)
    if family == "upper":
        # not vacuous: failures, and rows extending after a failure in the same walk
        fails = sum(1 for r in range(1, n) if flags[r] and (r % S) % 128 != 0)
        assert fails > 0 and fails < n - 1 - (n - 1) // 128
        wrong, wflags = two_phase_supernodes(E, nzu, S, restart_on_failure=True)
        assert wrong != o["nsup"] or not np.array_equal(wflags, o["sup_flag"])


def test_planted_restart_variant_disagrees_on_the_hand_case():
    D = rows_to_dense(3, [[0, 1, 2], [1, 2], [1, 2]])
    E = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)
    nzu = np.triu(E, 1).sum(axis=1)
    assert two_phase_supernodes(E, nzu, 128)[0] == 3
    assert two_phase_supernodes(E, nzu, 128, restart_on_failure=True)[0] == 2


def test_supernodes_reference_dataset_two_phase():
    """The reference's 11,948-row input at S = 128 and 1024: the oracle's count and flags equal the
    two-phase restatement run on the oracle's own F (F itself is pinned by 980,136 above)."""
    ds = load_ds()
    for S in (128, 1024):
        o = oracle.symbfact(ds, chunk=S, structure=True)
        n = ds.n
        ptr, cols = o["F_rowptr"], o["F_cols"]

        class RowsF:  # (r, s) in F by binary search in the sorted F row r
            def __getitem__(self, rs):
                r, s = rs
                row = cols[ptr[r]:ptr[r + 1]]
                k = np.searchsorted(row, s)
                return k < len(row) and row[k] == s
        count, flags = two_phase_supernodes(RowsF(), o["nzu"], S)
        assert o["nsup"] == count, S
        assert np.array_equal(o["sup_flag"], flags)


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
