"""GPU parity: the HIP engine against the oracle, bit-exact, through the C ABI."""
import os
import subprocess
import sys

import numpy as np
import pytest

import gsofatpds_amd as g
import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def load_ds():
    z = np.load(os.path.join(GOLDEN, "ds_11948.npz"))
    return g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])


def check(p, S=128, structure=True, label="", **opt):
    say(f"[parity] {label} n={p.n} nnz={p.nnz} S={S} structure={structure} opt={opt}")
    o = oracle.symbfact(p, chunk=S, structure=structure)
    with g.Plan(p, chunk_size=S, structure=int(structure), **opt) as plan:
        plan.run()
        r = plan.result()
        assert r.fill_count == o["fill"]
        assert r.nnz_F == o["nnz_F"]
        assert r.edge_checks == o["W"]
        assert r.reach == o["R"]
        assert r.supernodes_counted == o["nsup"]
        assert r.supernodes == o["nsup"] + 1
        assert np.array_equal(plan.nzu(), o["nzu"])
        assert np.array_equal(plan.supernode_flags(), o["sup_flag"])
        if structure:
            up, ui, lp, li = plan.structure()
            ou, oui, ol, oli = oracle.split_structure(p.n, o["F_rowptr"], o["F_cols"])
            assert np.array_equal(up, ou) and np.array_equal(ui, oui)
            assert np.array_equal(lp, ol) and np.array_equal(li, oli)
    return r


def test_device_is_gfx950():
    info = g.device_query(0)
    say(f"[device] {info}")
    assert info["arch"].startswith("gfx950")


def test_smallest_patterns():
    check(g.Pattern.from_dense(np.ones((1, 1), bool)), label="1x1")
    check(g.Pattern.from_dense(np.eye(2, dtype=bool)), label="2x2 diag")
    check(g.Pattern.from_dense(np.ones((3, 3), bool)), label="3x3 dense")


def test_grid_4096():
    for S in (128, 64, 256):
        check(g.generate_workload(g.WL_GRID2D_4096), S=S, label="grid2d_4096")


def test_reference_dataset():
    # structurally symmetric: traversed from A only by default, from A and A^T with symmetric=-1
    r = check(load_ds(), S=128, label="ds_11948")
    assert r.fill_count == 980136
    check(load_ds(), S=1024, structure=False, label="ds_11948")
    check(load_ds(), S=128, label="ds_11948 both directions", symmetric=-1)


@pytest.mark.parametrize("n,density,seed", [(7, 0.3, 1), (100, 0.05, 2), (257, 0.02, 3), (1000, 0.004, 4),
                                            (3000, 0.0015, 5), (3000, 0.003, 6)])
def test_random_unsymmetric(n, density, seed):
    rng = np.random.default_rng(seed)
    D = rng.random((n, n)) < density
    np.fill_diagonal(D, rng.random(n) < 0.9)  # some diagonals missing
    D[rng.integers(0, n, 3)] = False          # some empty rows
    check(g.Pattern.from_dense(D), label=f"random({n},{density})")


def test_long_rows_take_the_csr_tail():
    rng = np.random.default_rng(9)
    n = 2000
    D = rng.random((n, n)) < 0.002
    np.fill_diagonal(D, True)
    D[rng.integers(0, n, 20)] |= rng.random((20, n)) < 0.05  # rows of ~100 entries
    D[:, 1500] |= rng.random(n) < 0.1                         # a long column (long A^T row)
    check(g.Pattern.from_dense(D), label="long rows")


def test_duplicates_and_unsorted_rows():
    """Duplicate column ids are counted in nzU as the reference does; F and fills are sets."""
    rng = np.random.default_rng(4)
    n = 600
    D = rng.random((n, n)) < 0.01
    np.fill_diagonal(D, True)
    p = g.Pattern.from_dense(D)
    rows = []
    for v in range(n):
        r = list(p.adj[p.beg[v]:p.end[v]])
        if v % 7 == 0 and r:
            r += r[-2:]  # duplicates, including upper entries
        rng.shuffle(r)
        rows.append(np.array(r, np.int32))
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    q = g.Pattern.from_rowptr(ptr, np.concatenate(rows))
    check(q, label="duplicates")


def test_noncontiguous_rows():
    """Rows given by beg/end in arbitrary order inside adj, as in the reference's dataset."""
    p = g.generate_workload(g.WL_GRID2D_4096)
    order = np.random.default_rng(1).permutation(p.n)
    adj, beg, end, pos = [], np.zeros(p.n, np.int32), np.zeros(p.n, np.int32), 0
    for v in order:
        r = p.adj[p.beg[v]:p.end[v]]
        beg[v], end[v] = pos, pos + len(r)
        adj.append(r)
        pos += len(r)
    check(g.Pattern(p.n, beg, end, np.concatenate(adj)), label="noncontiguous")


@pytest.mark.parametrize("member", [6, 7, 11])
def test_batch_members(member):
    p = g.generate_workload(g.WL_BATCH256, member)
    check(p, structure=p.n < 3000, label=f"batch member {member}")


@pytest.mark.parametrize("chain_block", [1, 3, 64, 4096])
def test_chaining_blocks_grid3d_small(chain_block):
    """ND 3-D stencil (separators chain along grid lines) with several block sizes."""
    p = g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=14, ny=14, nz=14, order=g.ORDER_ND,
                   drop_side=g.DROP_UPPER, drop_ppm=150000, seed=77)
    check(p, S=128, label=f"grid3d 14^3 chain {chain_block}", chain_block=chain_block)


@pytest.mark.parametrize("chain_block", [1, 3, 64])
def test_paired_directions_forced(chain_block):
    """Workgroup-path items running both directions of a block (the other direction's reach set
    decides s-1 in R_s before any expansion), forced on structured and random patterns."""
    check(g.generate_workload(g.WL_GRID2D_4096), label=f"grid2d_4096 paired chain {chain_block}",
          pair_directions=1, chain_block=chain_block)
    rng = np.random.default_rng(21)
    n = 3000
    D = rng.random((n, n)) < 0.002
    np.fill_diagonal(D, rng.random(n) < 0.9)
    D[rng.integers(0, n, 10)] |= rng.random((10, n)) < 0.03  # long rows (CSR tails)
    D[:, 2500] |= rng.random(n) < 0.05                        # a long column
    check(g.Pattern.from_dense(D), label=f"random long rows paired chain {chain_block}",
          pair_directions=1, chain_block=chain_block)


@pytest.mark.parametrize("pair", [0, 1])
def test_paired_directions_reference_dataset(pair):
    # both directions (a mirrored matrix has one direction and no pairs)
    check(load_ds(), S=128, label=f"ds_11948 pair {pair}", pair_directions=pair, symmetric=-1)


@pytest.mark.parametrize("member", [7, 11, 1])
def test_erdos_renyi_members_pair_automatically(member):
    """Unstructured members (rows rarely hold their predecessor) take paired items by default."""
    p = g.generate_workload(g.WL_BATCH256, member)
    check(p, structure=p.n < 3000, label=f"batch member {member} (auto)")
    if p.n < 3000:
        check(p, structure=True, label=f"batch member {member} unpaired", pair_directions=0)


@pytest.mark.parametrize("case", ["grid2d_4096", "ds_11948", "random_long_rows_paired", "member_7", "grid3d_chain1"])
def test_bitmaps_in_hbm_forced(case):
    """The bitmap path with its bitmaps in HBM (L2 atomics, fenced barriers), forced at sizes the
    oracle checks: plain and paired items, long rows, chaining on and off, with the structure."""
    if case == "grid2d_4096":
        check(g.generate_workload(g.WL_GRID2D_4096), label="grid2d_4096 hbm", bitmap_in_hbm=1)
    elif case == "ds_11948":
        check(load_ds(), S=128, label="ds_11948 hbm", bitmap_in_hbm=1)
    elif case == "random_long_rows_paired":
        rng = np.random.default_rng(23)
        n = 3000
        D = rng.random((n, n)) < 0.002
        np.fill_diagonal(D, rng.random(n) < 0.9)
        D[rng.integers(0, n, 10)] |= rng.random((10, n)) < 0.03
        check(g.Pattern.from_dense(D), label="random long rows paired hbm", bitmap_in_hbm=1, pair_directions=1,
              chain_block=3)
    elif case == "member_7":
        p = g.generate_workload(g.WL_BATCH256, 7)
        check(p, structure=p.n < 3000, label="batch member 7 hbm", bitmap_in_hbm=1)
    else:
        check(g.generate_workload(g.WL_GRID2D_4096), label="grid2d_4096 hbm chain 1", bitmap_in_hbm=1,
              chain_block=1)


def test_large_n_takes_hbm_bitmaps():
    """n = 1,500,160 (a bitmap of 188 KB does not fit in LDS): 1,465 copies of one 1,024-row
    random block on the diagonal. Blocks never interact and align with the supernode chunks, so
    every total is the block's oracle total times the copies and nzU repeats the block's."""
    rng = np.random.default_rng(31)
    nb, copies = 1024, 1465
    D = rng.random((nb, nb)) < 6.0 / nb
    np.fill_diagonal(D, True)
    D &= ~(np.triu(rng.random((nb, nb)) < 0.1, 1))  # asymmetric: drop 10 % of the upper entries
    blk = g.Pattern.from_dense(D)
    o = oracle.symbfact(blk, chunk=128, structure=False)
    deg = blk.end - blk.beg
    beg = np.concatenate([[0], np.cumsum(np.tile(deg, copies))[:-1]]).astype(np.int32)
    end = (beg + np.tile(deg, copies)).astype(np.int32)
    rows_adj = np.concatenate([blk.adj[blk.beg[i]:blk.end[i]] for i in range(nb)]).astype(np.int64)
    adj = (np.tile(rows_adj, copies) + np.repeat(np.arange(copies, dtype=np.int64) * nb, len(rows_adj))).astype(np.int32)
    p = g.Pattern(nb * copies, beg, end, adj)
    say(f"[large n] n={p.n} nnz={p.nnz}")
    with g.Plan(p, chunk_size=128) as plan:
        plan.run()
        r = plan.result()
        say(f"[large n] {r.time_ms:.1f} ms, heavy sources {r.heavy_sources}, stages {plan.stage_times()}")
        assert (r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes_counted) == \
            (copies * o["fill"], copies * o["nnz_F"], copies * o["W"], copies * o["R"], copies * o["nsup"])
        assert np.array_equal(plan.nzu(), np.tile(o["nzu"], copies))


def test_tuning_configurations_are_identical():
    p = load_ds()
    base = None
    for opt in (dict(), dict(heavy_threads=256), dict(heavy_threads=1024), dict(queue_capacity=256),
                dict(light_wgs_per_cu=1, heavy_wgs_per_cu=1), dict(chain_block=1), dict(chain_block=7),
                dict(chain_block=1024, queue_capacity=256), dict(pair_directions=1, symmetric=-1),
                dict(pair_directions=1, chain_block=3, heavy_threads=256, symmetric=-1), dict(symmetric=-1),
                dict(cu_share=64), dict(cu_share=7, symmetric=-1)):
        say(f"[tuning] {opt}")
        with g.Plan(p, **opt) as plan:
            plan.run()
            r = plan.result()
            got = (r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes, plan.nzu().tobytes(),
                   plan.supernode_flags().tobytes())
        if base is None:
            base = got
        assert got == base


def test_repeated_runs_and_stream():
    p = g.generate_workload(g.WL_GRID2D_4096)
    s = g.Stream()
    with g.Plan(p, structure=1) as plan:
        plan.run(s).sync()
        first = plan.result().as_dict()
        st = plan.structure()
        for _ in range(3):
            plan.run(s).sync()
            r = plan.result().as_dict()
            first["time_ms"] = r["time_ms"]
            assert r == first
            st2 = plan.structure()
            assert all(np.array_equal(a, b) for a, b in zip(st, st2))


def test_batch_equals_plain():
    members = [6, 7, 11, 8, 2]
    pats = [g.generate_workload(g.WL_BATCH256, m) for m in members]
    say(f"[batch] members {members} n={[p.n for p in pats]}")
    b = g.Batch(pats)
    for rep in range(2):
        b.run().sync()
        for i, p in enumerate(pats):
            with g.Plan(p) as plan:
                plan.run()
                r = plan.result()
                rb = b.result(i)
                for k in ("fill_count", "nnz_F", "edge_checks", "reach", "supernodes"):
                    assert getattr(rb, k) == getattr(r, k), (rep, i, k)
                assert np.array_equal(b.nzu(i), plan.nzu())
                assert np.array_equal(b.supernode_flags(i), plan.supernode_flags())
    b.close()


def test_autotuned_batch_and_plan_are_identical():
    """The autotuner only picks launch shapes: results equal the untuned ones."""
    members = [6, 7, 11, 8, 2, 1]
    pats = [g.generate_workload(g.WL_BATCH256, m) for m in members]
    keys = ("fill_count", "nnz_F", "edge_checks", "reach", "supernodes")
    res = []
    for opt in (dict(), dict(autotune=1)):
        b = g.Batch(pats, **opt)
        say(f"[autotune] batch {opt}: {b.config()}")
        b.run().sync()
        res.append([(tuple(getattr(b.result(i), k) for k in keys), b.nzu(i).tobytes(), b.supernode_flags(i).tobytes())
                    for i in range(len(pats))])
        b.close()
    assert res[0] == res[1]
    check(load_ds(), S=128, label="ds_11948 autotuned", autotune=1)


def test_entry_and_cli_on_reference_files(tmp_path):
    ds = load_ds()
    files = [str(tmp_path / f) for f in ("colbeg_11948.dat", "colend_11948.dat", "csr_11948.dat")]
    g.write_ascii(ds, *files)
    argv = ["gsofa", *files, "1", "128", "0", "0", "1", "1", "4096"]
    fill, ms, nsup, teps = g.symbfact_min_id(argv)
    o = oracle.symbfact(ds, chunk=128)
    assert fill == 980136 and nsup == o["nsup"] + 1 and teps == o["W"] and ms > 0
    r = subprocess.run([g.CLI_PATH, *argv[1:]], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Final Number of fill-ins detected: 980136" in r.stdout
    assert f"Final Number of Supernodes detected: {o['nsup'] + 1}" in r.stdout
    csv = (tmp_path / "Vary_percent_blocksi_pre2.csv").read_text().splitlines()
    assert csv[0] == "vert_count;N_gpus;p0;p1;#blocks_src;min_time;max_time"
    assert csv[1].startswith(files[1] + ";1;0;0;1;")  # argv[2];argv[4];argv[6];argv[7];argv[8]


@pytest.mark.parametrize("case", ["ds_11948", "grid2d_4096"])
def test_supernodal_export(case):
    """The supernode partition and panels of a plan against the host export of the oracle's F and
    flags (the host export itself is pinned against the dense definition on the CPU)."""
    p = load_ds() if case == "ds_11948" else g.generate_workload(g.WL_GRID2D_4096)
    say(f"[supernodal] {case} n={p.n}")
    o = oracle.symbfact(p, chunk=128, structure=True)
    ref = g.supernodal_from_structure(o["sup_flag"], *oracle.split_structure_fast(p.n, o["F_rowptr"], o["F_cols"]))
    with g.Plan(p, chunk_size=128, structure=1) as plan:
        plan.run()
        xsup, supno = plan.supernode_partition()
        got = plan.supernodal_structure()
        assert len(xsup) - 1 == plan.result().supernodes_counted == o["nsup"]
    assert np.array_equal(xsup, ref[0])
    assert np.array_equal(supno, np.repeat(np.arange(len(xsup) - 1), np.diff(xsup)))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ib,vb", [(4, 4), (8, 8), (8, 4)])
def test_binary_files_through_the_engine(tmp_path, ib, vb):
    """The reference's dataset written as binary CSR (gSoFa_SOAComparison/graph.hpp layout) and read
    back runs through the engine: 980,136 fill-ins and the oracle's nzU and supernode flags."""
    ds = load_ds()
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    g.write_binary(ds, beg, csr, ib, vb)
    p = g.read_binary(beg, csr, ib, vb)
    say(f"[binary] widths {ib}/{vb} n={p.n} nnz={p.nnz}")
    o = oracle.symbfact(ds, chunk=128)
    with g.Plan(p, chunk_size=128) as plan:
        plan.run()
        r = plan.result()
        assert r.fill_count == 980136 == o["fill"]
        assert r.supernodes_counted == o["nsup"]
        assert np.array_equal(plan.nzu(), o["nzu"])
        assert np.array_equal(plan.supernode_flags(), o["sup_flag"])


def test_batch256_equals_plain_and_oracle():
    """The full 256-member batch (n log-uniform in [2,000, 200,000], half ND stencils, half
    Erdos-Renyi in RCM order) in one captured graph at the bench's tuned shape, replayed twice:
    every member's totals, nzU and supernode flags equal a plain plan's, and members below 12,000
    rows equal the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as ex:
        pats = list(ex.map(lambda k: g.generate_workload(g.WL_BATCH256, k), range(256)))
    say(f"[batch256] n total {sum(p.n for p in pats)}")
    keys = ("fill_count", "nnz_F", "edge_checks", "reach", "supernodes_counted")
    b = g.Batch(pats, heavy_threads=256, heavy_wgs_per_cu=4, chain_block=768)
    try:
        b.run().sync()
        first = [tuple(getattr(b.result(i), k) for k in keys) for i in range(256)]
        b.run().sync()
        res = [tuple(getattr(b.result(i), k) for k in keys) for i in range(256)]
        assert res == first  # replays agree
        nzu = [b.nzu(i) for i in range(256)]
        flags = [b.supernode_flags(i) for i in range(256)]
    finally:
        b.close()
    say("[batch256] batch done; plain plans and oracle")
    checked = 0
    for i, p in enumerate(pats):
        with g.Plan(p) as plan:
            plan.run()
            r = plan.result()
            assert tuple(getattr(r, k) for k in keys) == res[i], i
            assert np.array_equal(plan.nzu(), nzu[i]), i
            assert np.array_equal(plan.supernode_flags(), flags[i]), i
        if p.n < 12000:
            o = oracle.symbfact(p, chunk=128)
            assert res[i] == (o["fill"], o["nnz_F"], o["W"], o["R"], o["nsup"]), i
            assert np.array_equal(nzu[i], o["nzu"]) and np.array_equal(flags[i], o["sup_flag"]), i
            checked += 1
    say(f"[batch256] all members equal plain; {checked} checked against the oracle")
    assert checked >= 90


def test_results_before_any_run_are_a_state_error():
    p = g.generate_workload(g.WL_GRID2D_4096)
    with g.Plan(p) as plan:
        for call in (plan.result, plan.nzu, plan.supernode_flags, plan.stage_times, plan.sync):
            with pytest.raises(g.GsofaError) as e:
                call()
            assert e.value.code == g.ERR_STATE
    b = g.Batch([p, p])
    try:
        with pytest.raises(g.GsofaError) as e:
            b.result(0)
        assert e.value.code == g.ERR_STATE
        b.run().sync()
        assert b.result(1).fill_count == b.result(0).fill_count
    finally:
        b.close()


def test_stage_timing_switch():
    """Runs without stage events give the same results; their stage times are a state error
    until a run records them again."""
    p = g.generate_workload(g.WL_GRID2D_4096)
    with g.Plan(p, structure=1) as plan:
        plan.run()
        want = (plan.result().fill_count, g.fnv1a_array(plan.nzu()), g.fnv1a_array(plan.supernode_flags()))
        assert plan.stage_times()["heavy_ms"] > 0
        plan.stage_timing(False).run()
        assert (plan.result().fill_count, g.fnv1a_array(plan.nzu()), g.fnv1a_array(plan.supernode_flags())) == want
        with pytest.raises(g.GsofaError) as e:
            plan.stage_times()
        assert e.value.code == g.ERR_STATE
        plan.stage_timing(True).run()
        assert plan.result().fill_count == want[0] and plan.stage_times()["heavy_ms"] > 0


def _large(name):
    import hashlib
    import json
    path = os.path.join(GOLDEN, "golden_large.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    if name not in data:
        pytest.skip(f"no committed oracle values for {name}")
    return data[name], hashlib


@pytest.mark.parametrize("name,wl", [("grid3d_110592", g.WL_GRID3D_110592), ("grid2d_659033", g.WL_GRID2D_659033)])
def test_full_size_against_committed_oracle_values(name, wl):
    gd, hashlib = _large(name)
    p = g.generate_workload(wl)
    say(f"[full size] {name} n={p.n} nnz={p.nnz}")
    with g.Plan(p, chunk_size=128) as plan:
        plan.run()
        r = plan.result()
        o = gd["oracle_S128"]
        assert (r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes_counted) == \
            (o["fill"], o["nnz_F"], o["W"], o["R"], o["nsup"])
        assert hashlib.sha256(plan.nzu().astype(np.int32).tobytes()).hexdigest() == gd["nzu_sha256_S128"]
        assert hashlib.sha256(plan.supernode_flags().tobytes()).hexdigest() == gd["sup_flag_sha256_S128"]
        say(f"[full size] {name}: {r.time_ms:.2f} ms, heavy sources {r.heavy_sources}, stages {plan.stage_times()}")
    if name == "grid2d_659033":  # the same with the bitmaps in HBM
        with g.Plan(p, chunk_size=128, bitmap_in_hbm=1) as plan:
            plan.run()
            r = plan.result()
            assert (r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes_counted) == \
                (o["fill"], o["nnz_F"], o["W"], o["R"], o["nsup"])
            assert hashlib.sha256(plan.nzu().astype(np.int32).tobytes()).hexdigest() == gd["nzu_sha256_S128"]
            say(f"[full size] {name} hbm bitmaps: {r.time_ms:.2f} ms, stages {plan.stage_times()}")


@pytest.mark.parametrize("name,wl", [("grid3d_110592", g.WL_GRID3D_110592), ("grid2d_659033", g.WL_GRID2D_659033)])
def test_full_size_structure_equals_oracle(name, wl):
    """The explicit structure at full size, against digests of the oracle's F (strict U by rows,
    strict L by columns) committed by tests/golden/make_golden_large.py."""
    gd, _ = _large(name)
    if "U_rows_sha256" not in gd:
        pytest.skip(f"no committed structure digests for {name}")
    p = g.generate_workload(wl)
    say(f"[full size structure] {name} n={p.n}")
    with g.Plan(p, chunk_size=128, structure=1) as plan:
        plan.run()
        r = plan.result()
        up, ui, lp, li = plan.structure()
    assert (int(up[-1]), int(lp[-1])) == (gd["nnz_U"], gd["nnz_L"]) == (r.nnz_U, r.nnz_L)
    assert oracle.structure_digests(up, ui, lp, li) == (gd["U_rows_sha256"], gd["L_cols_sha256"])


@pytest.mark.parametrize("name,wl", [("grid3d_110592", g.WL_GRID3D_110592), ("grid2d_659033", g.WL_GRID2D_659033)])
def test_repeated_runs_equal_oracle(name, wl):
    """Which rows chain, escalate or take the workgroup path depends on timing (escalation flags
    and the wave path's sibling exchange are read while other waves write them); results must
    not. Several runs of one plan, each against the oracle's committed totals and digests."""
    gd, hashlib = _large(name)
    o = gd["oracle_S128"]
    p = g.generate_workload(wl)
    with g.Plan(p, chunk_size=128, structure=0) as plan:
        for k in range(6):
            say(f"[repeated] {name} run {k}")
            plan.run()
            r = plan.result()
            assert (r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes_counted) == \
                (o["fill"], o["nnz_F"], o["W"], o["R"], o["nsup"])
            assert hashlib.sha256(plan.nzu().astype(np.int32).tobytes()).hexdigest() == gd["nzu_sha256_S128"]
            assert hashlib.sha256(plan.supernode_flags().tobytes()).hexdigest() == gd["sup_flag_sha256_S128"]


@pytest.mark.parametrize("case", ["grid2d_nd_256", "grid3d_14_chain3", "ds_11948", "grid3d_110592"])
def test_block_cuts_change_nothing(case):
    """Blocks cut at large separator starts run as two workgroup-path items; results are those of
    uncut blocks (and of the oracle where it is cheap), with the structure."""
    if case == "grid2d_nd_256":
        p, opt, with_oracle = g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=256, ny=256, order=g.ORDER_ND,
                                         drop_side=g.DROP_LOWER, drop_ppm=100000, seed=5), {}, True
    elif case == "grid3d_14_chain3":
        p, opt, with_oracle = g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=14, ny=14, nz=14, order=g.ORDER_ND,
                                         seed=8), {"chain_block": 3}, True
    elif case == "ds_11948":
        p, opt, with_oracle = load_ds(), {}, True
    else:
        p, opt, with_oracle = g.generate_workload(g.WL_GRID3D_110592), {}, False
    say(f"[cuts] {case} n={p.n}")
    if with_oracle:
        check(p, label=f"{case} cut", **opt)
    out = []
    for split in (0, -1):
        with g.Plan(p, structure=1, split_blocks=split, **opt) as plan:
            plan.run()
            r = plan.result()
            up, ui, lp, li = plan.structure()
            out.append(((r.fill_count, r.nnz_F, r.edge_checks, r.reach, r.supernodes_counted),
                        oracle.structure_digests(up, ui, lp, li), g.fnv1a_array(plan.nzu()),
                        g.fnv1a_array(plan.supernode_flags())))
    assert out[0] == out[1]


def _symmetric_case(case):
    if case == "grid2d_nd_192":
        return g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=192, ny=192, order=g.ORDER_ND, seed=3), {}, True
    if case == "grid3d_14_chain3":
        return (g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=14, ny=14, nz=14, order=g.ORDER_ND, seed=8),
                {"chain_block": 3}, True)
    if case == "grid2d_natural_64":
        return g.generate(kind=g.GEN_STENCIL, dims=2, points=5, nx=64, ny=64, order=g.ORDER_NATURAL, seed=2), {}, True
    if case == "random_symmetric_2000":
        rng = np.random.default_rng(21)
        D = rng.random((2000, 2000)) < 0.002
        D = D | D.T
        np.fill_diagonal(D, rng.random(2000) < 0.8)  # some diagonals missing
        return g.Pattern.from_dense(D), {}, True
    # 110,592 rows, symmetric (the workload of that size drops entries of one triangle)
    return g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=48, ny=48, nz=48, order=g.ORDER_ND, seed=4), {}, False


@pytest.mark.parametrize("case", ["grid2d_nd_192", "grid3d_14_chain3", "grid2d_natural_64", "random_symmetric_2000",
                                  "grid3d_48_symmetric"])
def test_mirrored_symmetric_patterns(case):
    """A structurally symmetric pattern is traversed from A only, L mirrored from U: the oracle's
    results (where it is cheap) and exactly those of a run over both A and A^T, with less work
    executed."""
    p, opt, with_oracle = _symmetric_case(case)
    say(f"[mirror] {case} n={p.n}")
    if with_oracle:
        check(p, label=f"{case} mirrored", **opt)
    out, work = [], []
    for sym in (0, -1):
        with g.Plan(p, structure=1, symmetric=sym, **opt) as plan:
            plan.run()
            r = plan.result()
            up, ui, lp, li = plan.structure()
            out.append(((r.fill_count, r.nnz_F, r.nnz_L, r.nnz_U, r.edge_checks, r.reach, r.supernodes_counted),
                        oracle.structure_digests(up, ui, lp, li), g.fnv1a_array(plan.nzu()),
                        g.fnv1a_array(plan.supernode_flags())))
            w = plan.executed_work()
            work.append(w["light_edges"] + w["heavy_edges"])
    assert out[0] == out[1]
    # one traversal per source instead of two (the executed work varies a little with chaining)
    assert work[0] < 0.7 * work[1], work


def test_mirrored_and_unsymmetric_members_in_one_batch():
    """A set mixing symmetric members (mirrored) and unsymmetric ones (both directions, some
    paired): every member as in a set where nothing is mirrored, and as its own plan."""
    rng = np.random.default_rng(33)
    pats = []
    for k in range(7):
        if k % 2 == 0:
            pats.append(g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=40 + 13 * k, ny=37 + 11 * k,
                                   order=g.ORDER_ND, seed=k))
        else:
            n = 700 + 300 * k
            D = rng.random((n, n)) < 3.0 / n
            np.fill_diagonal(D, True)
            pats.append(g.Pattern.from_dense(D))
    say(f"[mirror batch] n={[p.n for p in pats]}")
    keys = ("fill_count", "nnz_F", "nnz_L", "nnz_U", "edge_checks", "reach", "supernodes")
    res = []
    for sym in (0, -1):
        b = g.Batch(pats, symmetric=sym)
        b.run().sync()
        res.append([tuple(getattr(b.result(i), k) for k in keys) + (g.fnv1a_array(b.nzu(i)),
                                                                    g.fnv1a_array(b.supernode_flags(i)))
                    for i in range(len(pats))])
        b.close()
    assert res[0] == res[1]
    for i, p in enumerate(pats):
        with g.Plan(p) as plan:
            plan.run()
            r = plan.result()
            assert tuple(getattr(r, k) for k in keys) + (g.fnv1a_array(plan.nzu()),
                                                          g.fnv1a_array(plan.supernode_flags())) == res[0][i]


def test_full_size_structure_properties():
    """Size-independent properties of the explicit structure at 659,033 rows: counts agree with
    the per-row counts, rows are sorted and duplicate-free, every entry of A is in F, and runs
    with different tuning agree."""
    p = g.generate_workload(g.WL_GRID2D_659033)
    with g.Plan(p, structure=1) as plan:
        plan.run()
        r = plan.result()
        up, ui, lp, li = plan.structure()
        nzu = plan.nzu()
    assert up[-1] == r.nnz_U and lp[-1] == r.nnz_L
    assert np.array_equal(np.diff(up), nzu)
    n = p.n
    rows = np.repeat(np.arange(n), np.diff(up))
    assert np.all(ui > rows)
    cols = np.repeat(np.arange(n), np.diff(lp))
    assert np.all(li > cols)
    # sorted and unique inside every row / column
    assert np.all(np.diff(ui.astype(np.int64))[rows[1:] == rows[:-1]] > 0)
    assert np.all(np.diff(li.astype(np.int64))[cols[1:] == cols[:-1]] > 0)
    # A is contained in F
    a_rows = np.repeat(np.arange(n), p.end - p.beg)
    upper = p.adj > a_rows
    ukeys = rows.astype(np.int64) * n + ui
    akeys = a_rows[upper].astype(np.int64) * n + p.adj[upper]
    assert np.all(np.isin(akeys, ukeys))
    lower = p.adj < a_rows
    lkeys = cols.astype(np.int64) * n + li  # (column j, row i) pairs of L
    akeys_l = p.adj[lower].astype(np.int64) * n + a_rows[lower]
    assert np.all(np.isin(akeys_l, lkeys))
    with g.Plan(p, structure=1, heavy_threads=512, queue_capacity=1024) as plan2:
        plan2.run()
        up2, ui2, lp2, li2 = plan2.structure()
    assert np.array_equal(up, up2) and np.array_equal(ui, ui2) and np.array_equal(lp, lp2) and np.array_equal(li, li2)


def _early_case(case):
    if case == "grid2d_600_unsym":  # both directions, one item per direction and block
        return g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=600, ny=600, order=g.ORDER_ND,
                          drop_side=g.DROP_LOWER, drop_ppm=100000, seed=12), {}
    if case == "grid2d_600_mirrored":  # direction 0 only, L mirrored from U
        return g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=600, ny=600, order=g.ORDER_ND, seed=13), {}
    # paired items: one item runs both directions of a block
    return (g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=600, ny=600, order=g.ORDER_ND,
                       drop_side=g.DROP_UPPER, drop_ppm=50000, seed=14), {"pair_directions": 1})


@pytest.mark.parametrize("case", ["grid2d_600_unsym", "grid2d_600_mirrored", "grid2d_600_paired"])
def test_early_top_launch_changes_nothing(case):
    """The top separator's items run in an early launch on a stream of their own, beside the wave
    path; the results (totals, structure, counts, supernode flags) are those of a run without it,
    over several runs of one plan."""
    p, opt = _early_case(case)
    say(f"[early top] {case} n={p.n}")
    out = []
    for early in (1, -1):
        with g.Plan(p, structure=1, early_top=early, **opt) as plan:
            for k in range(3 if early == 1 else 1):
                plan.run()
                r = plan.result()
                up, ui, lp, li = plan.structure()
                out.append(((r.fill_count, r.nnz_F, r.nnz_L, r.nnz_U, r.edge_checks, r.reach, r.supernodes_counted),
                            oracle.structure_digests(up, ui, lp, li), g.fnv1a_array(plan.nzu()),
                            g.fnv1a_array(plan.supernode_flags())))
                st, et = plan.stage_times(), plan.early_top()
                say(f"[early top] {case} early={early} run {k}: {r.time_ms:.2f} ms, {st}, {et}")
                if early == 1:
                    assert et["items"] > 0 and et["ms"] > 0 and et["edges"] > 0, et
                else:
                    assert et["items"] == 0 and et["edges"] == 0, et
    assert all(o == out[-1] for o in out), [o[0] for o in out]
    # counting only: the counts, flags and totals of the runs above
    with g.Plan(p, structure=0, early_top=1, **opt) as plan:
        plan.run()
        r = plan.result()
        assert plan.early_top()["items"] > 0
        assert ((r.fill_count, r.nnz_F, r.nnz_L, r.nnz_U, r.edge_checks, r.reach, r.supernodes_counted),
                g.fnv1a_array(plan.nzu()), g.fnv1a_array(plan.supernode_flags())) == \
            (out[-1][0], out[-1][2], out[-1][3])


def test_early_top_launch_in_a_batch_graph():
    """A batch whose largest member holds the set's majority takes the early launch inside its
    captured graph (a fork to the internal stream and a join); every replay equals plain plans
    without it."""
    big, _ = _early_case("grid2d_600_unsym")
    pats = [big] + [g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=60 + 7 * k, ny=50 + 5 * k,
                               order=g.ORDER_ND, seed=k) for k in range(4)]
    keys = ("fill_count", "nnz_F", "nnz_L", "nnz_U", "edge_checks", "reach", "supernodes")
    want = []
    for p in pats:
        with g.Plan(p, early_top=-1) as plan:
            plan.run()
            r = plan.result()
            want.append(tuple(getattr(r, k) for k in keys) + (g.fnv1a_array(plan.nzu()),
                                                              g.fnv1a_array(plan.supernode_flags())))
    b = g.Batch(pats, early_top=1)
    try:
        for rep in range(3):
            say(f"[early top batch] replay {rep}")
            b.run().sync()
            got = [tuple(getattr(b.result(i), k) for k in keys) + (g.fnv1a_array(b.nzu(i)),
                                                                   g.fnv1a_array(b.supernode_flags(i)))
                   for i in range(len(pats))]
            assert got == want
            g.device_synchronize()
    finally:
        b.close()


@pytest.mark.parametrize("case", ["grid2d_nd_256", "grid3d_14_chain3", "ds_11948", "grid2d_nd_192_mirrored",
                                  "grid3d_24_paired"])
def test_pre_escalation_changes_nothing(case):
    """Rows whose elimination subtree exceeds the escalation limit skip the wave path; results (with
    the structure) are those of runs that attempt every row, and the oracle's where it is cheap."""
    if case == "grid2d_nd_256":
        p, opt = g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=256, ny=256, order=g.ORDER_ND,
                            drop_side=g.DROP_LOWER, drop_ppm=100000, seed=5), {}
    elif case == "grid3d_14_chain3":
        p, opt = g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=14, ny=14, nz=14, order=g.ORDER_ND,
                            seed=8), {"chain_block": 3}
    elif case == "ds_11948":
        p, opt = load_ds(), {}
    elif case == "grid2d_nd_192_mirrored":
        p, opt = g.generate(kind=g.GEN_STENCIL, dims=2, points=9, nx=192, ny=192, order=g.ORDER_ND, seed=3), {}
    else:
        p, opt = g.generate(kind=g.GEN_STENCIL, dims=3, points=7, nx=24, ny=24, nz=24, order=g.ORDER_ND,
                            drop_side=g.DROP_UPPER, drop_ppm=80000, seed=9), {"pair_directions": 1}
    say(f"[pre-escalation] {case} n={p.n}")
    check(p, label=f"{case} pre-escalation", **opt)
    out = []
    for pre in (0, -1):
        with g.Plan(p, structure=1, pre_escalate=pre, **opt) as plan:
            plan.run()
            r = plan.result()
            up, ui, lp, li = plan.structure()
            out.append(((r.fill_count, r.nnz_F, r.nnz_L, r.nnz_U, r.edge_checks, r.reach, r.supernodes_counted),
                        oracle.structure_digests(up, ui, lp, li), g.fnv1a_array(plan.nzu()),
                        g.fnv1a_array(plan.supernode_flags())))
    assert out[0] == out[1]


def test_pre_escalation_leaves_unsymmetric_bands_alone():
    """A lower band: direction 0 reaches every row below, direction 1 almost nothing, while every
    row's elimination subtree is huge. Pre-escalation must not send direction 1's rows to the
    workgroup path (the pattern is far from symmetric); results equal the oracle's."""
    n = 20000
    beg, adj = [], []
    for i in range(n):
        beg.append(len(adj))
        adj.extend(j for j in (i - 3, i - 2, i - 1, i) if j >= 0)
    beg = np.array(beg, np.int32)
    end = np.append(beg[1:], len(adj)).astype(np.int32)
    p = g.Pattern(n, beg, end, np.array(adj, np.int32))
    check(p, label="lower band", structure=False)
    hs = []
    for pre in (0, -1):
        with g.Plan(p, pre_escalate=pre) as plan:
            plan.run()
            hs.append(plan.result().heavy_sources)
    say(f"[band] workgroup-path traversals with / without pre-escalation: {hs}")
    assert hs[0] <= 1.05 * hs[1] + 64, hs
