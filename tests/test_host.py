"""Host-side checks without a GPU: the C ABI library loads and exports what include/gsofa.h
declares, generators, the reference's ASCII format, input validation and error codes."""
import os
import re
import subprocess

import numpy as np
import pytest

import gsofatpds_amd as g
from conftest import ROOT, gpu_available


def declared_symbols():
    with open(os.path.join(ROOT, "include", "gsofa.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(gsofa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = g.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_compiles_as_c():
    src = os.path.join(ROOT, "include", "gsofa.h")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_binding_structs_match_the_header(tmp_path):
    """The ctypes mirrors of gsofa_options, gsofa_result and gsofa_server_options have the C
    sizes and field offsets, and the defaults are the documented ones."""
    import ctypes
    prog = tmp_path / "sizes.c"
    fields = [f for f, _ in g.Options._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gsofa.h"', "int main(void) {",
             '  printf("%zu %zu %zu\\n", sizeof(gsofa_options), sizeof(gsofa_result), sizeof(gsofa_server_options));']
    lines += [f'  printf("%zu\\n", offsetof(gsofa_options, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    prog.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "sizes"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes, offsets = [int(x) for x in out[:3]], [int(x) for x in out[3:]]
    assert sizes == [ctypes.sizeof(g.Options), ctypes.sizeof(g.Result), ctypes.sizeof(g.ServerOptions)]
    assert offsets == [getattr(g.Options, f).offset for f in fields]
    o = g.default_options()
    assert (o.chunk_size, o.pair_directions, o.autotune, o.bitmap_in_hbm, o.structure) == (128, -1, 0, 0, 0)


def test_workload_sizes():
    p1 = g.generate_workload(g.WL_GRID2D_4096)
    assert p1.n == 4096 and 16000 < p1.nnz < 18000
    p2 = g.generate_workload(g.WL_GRID3D_110592)
    assert p2.n == 110592 and 650_000 < p2.nnz < 760_000


@pytest.mark.slow
def test_workload_659033():
    p3 = g.generate_workload(g.WL_GRID2D_659033)
    assert p3.n == 659033 and 5_500_000 < p3.nnz < 6_000_000


def test_generators_deterministic_and_well_formed():
    for wl, m in ((g.WL_GRID2D_4096, 0), (g.WL_BATCH256, 6), (g.WL_BATCH256, 7), (g.WL_STREAM, 3)):
        a = g.generate_workload(wl, m)
        b = g.generate_workload(wl, m)
        assert np.array_equal(a.adj, b.adj) and np.array_equal(a.beg, b.beg)
        for v in range(a.n):
            row = a.adj[a.beg[v]:a.end[v]]
            assert v in row  # diagonal always present
            assert np.all(np.diff(row) > 0)  # sorted, no duplicates
            assert row.min() >= 0 and row.max() < a.n


def test_batch_member_sizes():
    ns = [g.generate_workload(g.WL_BATCH256, k).n for k in range(0, 256, 17)]
    assert all(2000 <= n <= 200000 for n in ns)


def test_nd_ordering_is_a_permutation_with_separators_last():
    # 2-D 5-point 20x20 ND: the last vertex belongs to the top separator, a full grid line
    p = g.generate(kind=g.GEN_STENCIL, dims=2, points=5, nx=20, ny=20, order=g.ORDER_ND, nd_leaf=16)
    assert p.n == 400
    rows = {v: set(p.adj[p.beg[v]:p.end[v]]) for v in range(p.n)}
    # symmetric pattern (no drops)
    assert all(v in rows[u] for v in rows for u in rows[v])
    # the top separator (20 vertices, ids 380..399) splits the rest into two halves with no edges
    sep = set(range(380, 400))
    rest = [v for v in range(p.n) if v not in sep]
    comp = {}
    for s in rest:
        if s in comp:
            continue
        stack = [s]
        comp[s] = s
        while stack:
            v = stack.pop()
            for u in rows[v]:
                if u not in sep and u not in comp:
                    comp[u] = s
                    stack.append(u)
    assert len(set(comp.values())) == 2


def test_ascii_round_trip(tmp_path):
    p = g.generate_workload(g.WL_GRID2D_4096)
    files = [str(tmp_path / f) for f in ("beg.dat", "end.dat", "csr.dat")]
    g.write_ascii(p, *files)
    q = g.read_ascii(*files)
    assert q.n == p.n and np.array_equal(q.adj, p.adj) and np.array_equal(q.beg, p.beg)


def test_ascii_reads_reference_dataset_if_present():
    d = "/root/reference/SpaceOptimization/dataset"
    if not os.path.isdir(d):
        pytest.skip("reference dataset not mounted")
    p = g.read_ascii(f"{d}/colbeg_11948.dat", f"{d}/colend_11948.dat", f"{d}/csr_11948.dat")
    z = np.load(os.path.join(ROOT, "tests", "golden", "ds_11948.npz"))
    assert p.n == 11948 and np.array_equal(p.adj, z["adj"]) and np.array_equal(p.beg, z["beg"])


def test_ascii_errors(tmp_path):
    with pytest.raises(g.GsofaError) as e:
        g.read_ascii(str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c"))
    assert e.value.code == g.ERR_IO
    (tmp_path / "c").write_text("5 3\n0 1 2 0 9\n")
    (tmp_path / "a").write_text("0 2 4\n")
    (tmp_path / "b").write_text("2 4 5\n")
    with pytest.raises(g.GsofaError) as e:
        g.read_ascii(str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c"))
    assert e.value.code == g.ERR_BAD_CSR  # column id 9 >= n = 3


@pytest.mark.parametrize("ib,vb", [(8, 8), (4, 4), (8, 4)])
def test_binary_round_trip(tmp_path, ib, vb):
    """Reference binary CSR pair: n+1 offsets, then the rows back to back (rows given anywhere
    in adj come out contiguous and in row order)."""
    p = g.generate_workload(g.WL_GRID2D_4096)
    order = np.random.default_rng(3).permutation(p.n)
    adj, beg, end, pos = [], np.zeros(p.n, np.int32), np.zeros(p.n, np.int32), 0
    for v in order:  # non-contiguous input rows
        r = p.adj[p.beg[v]:p.end[v]]
        beg[v], end[v] = pos, pos + len(r)
        adj.append(r)
        pos += len(r)
    q = g.Pattern(p.n, beg, end, np.concatenate(adj))
    b, c = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    g.write_binary(q, b, c, ib, vb)
    assert os.path.getsize(b) == (p.n + 1) * ib and os.path.getsize(c) == p.nnz * vb
    off = np.fromfile(b, dtype=np.int64 if ib == 8 else np.int32)
    assert off[0] == 0 and off[-1] == p.nnz and np.all(np.diff(off) == p.end - p.beg)
    r = g.read_binary(b, c, ib, vb)
    assert r.n == p.n and np.array_equal(r.beg, p.beg) and np.array_equal(r.end, p.end)
    assert np.array_equal(r.adj, p.adj)


def test_binary_errors(tmp_path):
    b, c = tmp_path / "beg.bin", tmp_path / "csr.bin"
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c))
    assert e.value.code == g.ERR_IO                      # missing file
    np.array([0, 2, 4, 3], np.int64).tofile(b)           # decreasing offsets
    np.array([0, 1, 1, 2], np.int64).tofile(c)
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c))
    assert e.value.code == g.ERR_IO
    np.array([0, 2, 4, 6], np.int64).tofile(b)           # csr holds 4 ids, offsets need 6
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c))
    assert e.value.code == g.ERR_IO
    np.array([0, 1, 1, 2, 2, 5], np.int64).tofile(c)     # id 5 >= n = 3
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c))
    assert e.value.code == g.ERR_IO
    b.write_bytes(b"\x00" * 13)                          # not a whole number of offsets
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c))
    assert e.value.code == g.ERR_IO
    with pytest.raises(g.GsofaError) as e:
        g.read_binary(str(b), str(c), 2, 8)
    assert e.value.code == g.ERR_INVALID_ARG


def bad_patterns():
    ok = g.Pattern.from_dense(np.eye(4, dtype=bool))
    col = g.Pattern(4, ok.beg.copy(), ok.end.copy(), ok.adj.copy())
    col.adj[2] = 7
    ext = g.Pattern(4, ok.beg.copy(), ok.end.copy(), ok.adj.copy())
    ext.beg[1], ext.end[1] = 3, 2
    past = g.Pattern(4, ok.beg.copy(), ok.end.copy(), ok.adj.copy())
    past.end[3] = 9
    neg = g.Pattern(4, ok.beg.copy(), ok.end.copy(), ok.adj.copy())
    neg.adj[0] = -1
    return [col, ext, past, neg]


@pytest.mark.parametrize("idx", range(4))
def test_malformed_csr_is_rejected_before_any_device_work(idx):
    p = bad_patterns()[idx]
    with pytest.raises(g.GsofaError) as e:
        g.Plan(p)
    assert e.value.code == g.ERR_BAD_CSR


def test_entry_argument_count():
    with pytest.raises(g.GsofaError) as e:
        g.symbfact_min_id(["x", "a", "b"])
    assert e.value.code == g.ERR_INVALID_ARG


def test_entry_missing_file(tmp_path):
    argv = ["gsofa", str(tmp_path / "b"), str(tmp_path / "e"), str(tmp_path / "c"), "1", "128", "0", "0", "1", "1",
            "4096"]
    with pytest.raises(g.GsofaError) as e:
        g.symbfact_min_id(argv)
    assert e.value.code == g.ERR_IO


def test_no_device_is_an_error_code_not_a_fallback():
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(g.GsofaError) as e:
        g.Plan(g.Pattern.from_dense(np.eye(8, dtype=bool)))
    assert e.value.code == g.ERR_NO_DEVICE


def test_cli_usage():
    r = subprocess.run([g.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


# ---- conversion: Matrix Market and the SuperLU_DIST-style reordering -------------------------

def dense_of(p):
    D = np.zeros((p.n, p.n), np.int32)
    for v in range(p.n):
        np.add.at(D[v], p.adj[p.beg[v]:p.end[v]], 1)
    return D


def write_mtx(path, D, field="pattern", symmetry="general", lower_only=False):
    rows, cols = np.nonzero(np.tril(D) if lower_only else D)
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n% a comment\n")
        f.write(f"{D.shape[0]} {D.shape[1]} {len(rows)}\n")
        for i, j in zip(rows, cols):
            val = {"pattern": "", "real": " 1.5e-3", "integer": " 7", "complex": " 1.0 -2.0"}[field]
            f.write(f"{i + 1} {j + 1}{val}\n")


@pytest.mark.parametrize("field", ["pattern", "real", "integer", "complex"])
def test_matrix_market_general(tmp_path, field):
    rng = np.random.default_rng(1)
    D = (rng.random((50, 50)) < 0.08).astype(np.int32)
    write_mtx(tmp_path / "a.mtx", D, field)
    p = g.read_matrix_market(str(tmp_path / "a.mtx"))
    assert p.n == 50 and np.array_equal(dense_of(p), D)


def test_matrix_market_symmetric_expands(tmp_path):
    rng = np.random.default_rng(2)
    D = (rng.random((40, 40)) < 0.1).astype(np.int32)
    D = ((D + D.T) > 0).astype(np.int32)
    write_mtx(tmp_path / "s.mtx", D, "real", "symmetric", lower_only=True)
    p = g.read_matrix_market(str(tmp_path / "s.mtx"))
    assert np.array_equal(dense_of(p), D)


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", g.ERR_IO),
    ("2 2 1\n1 1\n", g.ERR_IO),
    ("%%MatrixMarket matrix coordinate pattern general\n2 3 1\n1 1\n", g.ERR_IO),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n3 1\n", g.ERR_BAD_CSR),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n", g.ERR_IO),
])
def test_matrix_market_errors(tmp_path, text, code):
    (tmp_path / "e.mtx").write_text(text)
    with pytest.raises(g.GsofaError) as e:
        g.read_matrix_market(str(tmp_path / "e.mtx"))
    assert e.value.code == code


def test_permute_is_superlu_dist_reordering():
    """B = Pc Pr A Pc^T with B(perm_c[perm_r[i]], perm_c[j]) = A(i, j): checked as dense matrix
    products with permutation matrices (P[perm[i], i] = 1)."""
    rng = np.random.default_rng(3)
    n = 60
    D = (rng.random((n, n)) < 0.07).astype(np.int64)
    p = g.Pattern.from_dense(D.astype(bool))
    pr, pc = rng.permutation(n).astype(np.int32), rng.permutation(n).astype(np.int32)
    Pr = np.zeros((n, n), np.int64)
    Pr[pr, np.arange(n)] = 1
    Pc = np.zeros((n, n), np.int64)
    Pc[pc, np.arange(n)] = 1
    B = dense_of(g.permute(p, pr, pc))
    assert np.array_equal(B, Pc @ Pr @ D @ Pc.T)
    assert np.array_equal(dense_of(g.permute(p, pr, None)), Pr @ D)
    assert np.array_equal(dense_of(g.permute(p, None, None)), D)
    inv = np.argsort(pc).astype(np.int32)  # Pc^T Pc = I: permuting back restores A
    assert np.array_equal(dense_of(g.permute(g.permute(p, None, pc), None, inv)), D)
    with pytest.raises(g.GsofaError) as e:
        g.permute(p, np.zeros(n, np.int32), None)
    assert e.value.code == g.ERR_INVALID_ARG


def test_convert_tool_reference_dataset(tmp_path):
    """The reference's input as a Matrix Market file through gsofa_convert with identity
    permutations gives the reference's ASCII triplet back; with a symmetric reordering the
    converted matrix is P A P^T."""
    import oracle
    z = np.load(os.path.join(ROOT, "tests", "golden", "ds_11948.npz"))
    ds = g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])
    rows = np.repeat(np.arange(ds.n), ds.end - ds.beg)
    cols = np.concatenate([ds.adj[b:e] for b, e in zip(ds.beg, ds.end)])
    with open(tmp_path / "ds.mtx", "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate pattern general\n{ds.n} {ds.n} {len(rows)}\n")
        f.write("".join(f"{i + 1} {j + 1}\n" for i, j in zip(rows, cols)))
    ident = tmp_path / "ident.txt"
    ident.write_text(" ".join(str(i + 1) for i in range(ds.n)))
    r = subprocess.run([g.CONVERT_PATH, str(tmp_path / "ds.mtx"), str(tmp_path / "out"), "--perm-r", str(ident),
                        "--perm-c", str(ident), "--one-based"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    q = g.read_ascii(*(str(tmp_path / f"out.{x}") for x in ("beg", "end", "csr")))
    assert q.n == ds.n and q.nnz == ds.nnz
    o = oracle.symbfact(q, chunk=128)
    assert o["fill"] == 980136
    perm = np.random.default_rng(4).permutation(ds.n)
    (tmp_path / "perm.txt").write_text(" ".join(map(str, perm)))
    r = subprocess.run([g.CONVERT_PATH, str(tmp_path / "ds.mtx"), str(tmp_path / "b"), "--perm-c",
                        str(tmp_path / "perm.txt"), "--binary", "8", "4"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    b = g.read_binary(str(tmp_path / "b.beg.bin"), str(tmp_path / "b.csr.bin"), 8, 4)
    ref = g.permute(ds, None, perm.astype(np.int32))
    for v in range(0, ds.n, 97):
        assert np.array_equal(np.sort(b.adj[b.beg[v]:b.end[v]]), np.sort(ref.adj[ref.beg[v]:ref.end[v]]))


def test_convert_tool_usage():
    r = subprocess.run([g.CONVERT_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
