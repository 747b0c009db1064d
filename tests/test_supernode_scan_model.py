"""Host model of k_supernodes' start search (gsofatpds_amd/csrc/kernels.hip): every workgroup marks
the supernode starts of its 256 rows and of the 128 rows before them as bits and finds a row's
start with word scans. Checked here, on random counts and window bits, against the row-by-row walk
of the reference's rule (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:1833, 1879,
1928 for the starts, :2034 and :2047 for pass-through rows). The window bits are random, so
this checks the search, not the oracle. CPU only."""
import random


def _is_start(nzu, chunk, r):
    return ((r % chunk) & 127) == 0 or nzu[r - 1] - nzu[r] != 1


def _walk(nzu, chunk, win):
    flags = []
    for r in range(len(nzu)):
        if _is_start(nzu, chunk, r):
            flags.append(1)
            continue
        s = r - 1
        while not _is_start(nzu, chunk, s):
            s -= 1
        t = r - s - 1
        flags.append(0 if (win[s][t >> 5] >> (t & 31)) & 1 else 1)
    return flags


def _bit_scan(nzu, chunk, win, back=128, rows=256):
    n = len(nzu)
    flags = [None] * n
    for r0 in range(0, n, rows):
        cnt = [nzu[row] if 0 <= row < n else 0 for row in range(r0 - back - 1, r0 + rows)]
        st = [0] * ((back + rows) // 32)
        for j in range(back + rows):
            row = r0 - back + j
            if 0 <= row < n and (((row % chunk) & 127) == 0 or cnt[j] - cnt[j + 1] != 1):
                st[j >> 5] |= 1 << (j & 31)
        for tid in range(rows):
            r, j = r0 + tid, back + tid
            if r >= n:
                continue
            if (st[j >> 5] >> (j & 31)) & 1:
                flags[r] = 1
                continue
            w = j >> 5
            bits = st[w] & ((1 << (j & 31)) - 1)
            while bits == 0 and w > 0:
                w -= 1
                bits = st[w]
            assert bits, "a start lies within 127 rows"
            s = r0 - back + w * 32 + bits.bit_length() - 1
            t = r - s - 1
            assert 0 <= t < 127
            flags[r] = 0 if (win[s][t >> 5] >> (t & 31)) & 1 else 1
    return flags


def test_bit_scan_equals_walk():
    rng = random.Random(1)
    for _ in range(200):
        n = rng.randint(1, 1200)
        chunk = rng.choice([1, 2, 7, 64, 100, 128, 129, 200, 256, 1000])
        nzu, v = [], rng.randint(0, 400)
        for _r in range(n):  # long runs of differences of 1 (long supernode candidates) and jumps
            nzu.append(v)
            v = v - 1 if rng.random() < 0.97 else rng.randint(0, 400)
        win = [[rng.getrandbits(32) if rng.random() < 0.5 else 0 for _ in range(4)] for _ in range(n)]
        assert _bit_scan(nzu, chunk, win) == _walk(nzu, chunk, win)


def test_a_planted_off_by_one_is_caught():
    """The same scan taking the start one row too early must disagree somewhere."""
    rng = random.Random(2)
    n, chunk = 900, 128
    nzu, v = [], 300
    for _r in range(n):
        nzu.append(v)
        v = v - 1 if rng.random() < 0.95 else rng.randint(0, 300)
    win = [[rng.getrandbits(32) for _ in range(4)] for _ in range(n)]
    good = _walk(nzu, chunk, win)

    def shifted(nzu, chunk, win):
        out = []
        for r in range(len(nzu)):
            if _is_start(nzu, chunk, r):
                out.append(1)
                continue
            s = r - 1
            while not _is_start(nzu, chunk, s):
                s -= 1
            t = r - s  # planted: one row off
            out.append(0 if t < 128 and (win[s][t >> 5] >> (t & 31)) & 1 else 1)
        return out
    assert shifted(nzu, chunk, win) != good
