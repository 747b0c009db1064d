"""Supernodal export on the host (no device): the partition and the per-supernode panels of
gsofa_supernodal_from_structure against a dense restatement from the definition, on F from
dense elimination and on the reference's 11,948-row input."""
import os

import numpy as np
import pytest

import gsofatpds_amd as g
import oracle
from conftest import GOLDEN


def dense_panels(E, flags):
    """The partition (a new supernode at row 0 and at every flagged row) and, per supernode
    [a, b), the rows i >= b holding F(i, j) for some j in [a, b) and the columns m >= b holding
    F(i, m) for some i in [a, b), from a dense F."""
    n = len(flags)
    xsup = [i for i in range(n) if i == 0 or flags[i]] + [n]
    lsub, usub = [], []
    for a, b in zip(xsup[:-1], xsup[1:]):
        lsub.append(np.nonzero(E[b:, a:b].any(axis=1))[0] + b)
        usub.append(np.nonzero(E[a:b, b:].any(axis=0))[0] + b)
    return np.array(xsup), lsub, usub


def strict_parts(E):
    n = E.shape[0]
    U = np.triu(E, 1)
    L = np.tril(E, -1)
    up = np.zeros(n + 1, np.int64)
    up[1:] = np.cumsum(U.sum(axis=1))
    ui = np.nonzero(U)[1].astype(np.int32)  # row-major: columns of each row, ascending
    lp = np.zeros(n + 1, np.int64)
    lp[1:] = np.cumsum(L.sum(axis=0))
    li = np.nonzero(L.T)[1].astype(np.int32)  # rows of each column, ascending
    return up, ui, lp, li


def check(E, flags):
    xsup, xl, ls, xu, us = g.supernodal_from_structure(flags, *strict_parts(E))
    exs, els, eus = dense_panels(E, flags)
    assert np.array_equal(xsup, exs)
    assert xl[0] == 0 and xu[0] == 0
    for k in range(len(exs) - 1):
        assert np.array_equal(ls[xl[k]:xl[k + 1]], els[k]), k
        assert np.array_equal(us[xu[k]:xu[k + 1]], eus[k]), k
    return xsup


@pytest.mark.parametrize("n,density,S,seed", [(60, 0.05, 128, 1), (150, 0.03, 64, 2), (256, 0.02, 256, 3),
                                              (200, 0.01, 128, 4)])
def test_panels_equal_dense_definition(n, density, S, seed):
    rng = np.random.default_rng(seed)
    D = rng.random((n, n)) < density
    np.fill_diagonal(D, True)
    E = oracle.dense_elimination(D)
    np.fill_diagonal(E, True)
    o = oracle.symbfact(g.Pattern.from_dense(D), chunk=S)
    xsup = check(E, o["sup_flag"])
    assert len(xsup) - 1 == o["nsup"]  # one supernode per increment of the count


def test_panels_single_supernode_and_all_singletons():
    n = 40
    E = np.ones((n, n), bool)
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    xsup = check(E, flags)  # one supernode: both panels empty
    assert list(xsup) == [0, n]
    xsup = check(E, np.ones(n, np.uint8))
    assert list(xsup) == list(range(n + 1))


def test_panels_reference_dataset():
    """The reference's 11,948-row input, S = 128: the oracle's F and flags through the host export
    against the dense definition."""
    z = np.load(os.path.join(GOLDEN, "ds_11948.npz"))
    p = g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])
    o = oracle.symbfact(p, chunk=128, structure=True)
    n = p.n
    E = np.zeros((n, n), bool)
    rows = np.repeat(np.arange(n), np.diff(o["F_rowptr"]))
    E[rows, o["F_cols"]] = True
    xsup = check(E, o["sup_flag"])
    assert len(xsup) - 1 == o["nsup"]


def test_bad_arguments():
    with pytest.raises(g.GsofaError):
        g.supernodal_from_structure(np.zeros(0, np.uint8), np.zeros(1, np.int64), [], np.zeros(1, np.int64), [])
