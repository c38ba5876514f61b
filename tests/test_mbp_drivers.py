"""Slice-time and fixed-tree MBP drivers on the CPU: the area tree of
Data::area_split (src/data_matrix.cc:299-437) and its split fit, the
proposals' masks and the sim_frac adaptation (src/mbp.cc:656-716,
src/utils.cc:619-623)."""
import math

import numpy as np
import pytest

from beep_amd import configs, mbp_drivers as md


def _two_clusters(A=12):
    """Areas 0..5 mixing strongly among themselves, 6..11 likewise, and
    weakly across."""
    spec, _ = configs.seir_areas(A, 1, 10, seed=3)
    diag = np.ones(A)
    to, val, rowptr = [], [], [0]
    for c in range(A):
        for j in range(A):
            if j != c:
                to.append(j)
                val.append(1.0 if (j < A // 2) == (c < A // 2) else 0.01)
        rowptr.append(len(to))
    spec = dict(spec, M={"diag": diag.tolist(), "rowptr": rowptr, "to": to, "val": val})
    return spec


def test_split_fit_dif_is_the_change_of_the_fit():
    rng = np.random.default_rng(1)
    G = rng.random((9, 9))
    areas = [7, 2, 5, 0, 8, 1]
    gr = rng.random(6) < 0.5
    for isel in range(6):
        before = md.split_fit(areas, gr, G)
        g2 = gr.copy()
        d = md.split_fit_dif(areas, g2, G, isel)
        assert g2[isel] != gr[isel] and (np.delete(g2, isel) == np.delete(gr, isel)).all()
        assert d == pytest.approx(md.split_fit(areas, g2, G) - before, rel=1e-12, abs=1e-12)


def test_split_fit_brute_force():
    G = np.arange(16.0).reshape(4, 4)
    gr = np.array([True, False, True, True])
    want = 0.0
    for i in range(4):
        for j in range(4):
            if i != j and G[j, i] != 0:
                want += G[j, i] if gr[i] == gr[j] else -G[j, i]
    assert md.split_fit([0, 1, 2, 3], gr, G) == want


def test_area_tree_partitions_and_leaves():
    spec, _ = configs.seir_areas(23, 1, 10, seed=5)
    tree = md.area_split(spec, np.random.default_rng(2))
    assert sorted(tree[0].arearef) == list(range(23))
    leaves = 0
    for n in tree:
        if len(n.arearef) > 1:
            a, b = (tree[c].arearef for c in n.child)
            assert sorted(a + b) == sorted(n.arearef) and not set(a) & set(b)
            # the swaps keep the halves' sizes of the first allocation
            assert len(a) == len(n.arearef) // 2 and len(b) == len(n.arearef) - len(n.arearef) // 2
        else:
            assert n.child == []
            leaves += 1
    assert leaves == 23 and len(tree) == 2 * 23 - 1


def test_area_split_finds_clusters_and_improves_the_fit():
    spec = _two_clusters()
    G = md.split_matrix(spec)
    tree = md.area_split(spec, np.random.default_rng(4))
    halves = sorted(sorted(tree[c].arearef) for c in tree[0].child)
    assert halves == [list(range(6)), list(range(6, 12))]
    # From a scrambled list the random-swap search keeps only improving swaps:
    # the split it returns fits at least as well as the first allocation (it
    # may stop at a local optimum: 50 failed swaps in a row end it).
    for seed in range(5):
        perm = np.random.default_rng(9 + seed).permutation(12).tolist()
        c1, c2 = md.split_in_two(perm, G, np.random.default_rng(7 + seed))
        gr = np.array([a in c1 for a in perm])
        assert md.split_fit(perm, gr, G) >= md.split_fit(perm, np.arange(12) < 6, G)
        assert md.split_fit(perm, gr, G) <= md.split_fit(list(range(12)), np.arange(12) < 6, G) + 1e-12


def test_normalised_mixing_rows():
    spec, _ = configs.seir_areas(9, 4, 10, seed=7)
    diag, rowptr, to, val, pop = md.normalised_mixing(spec)
    for c in range(9):
        a, b = rowptr[c], rowptr[c + 1]
        assert diag[c] * pop[c] + (val[a:b] * pop[to[a:b]]).sum() == pytest.approx(1.0, rel=1e-12)


def test_adaptation_closed_form():
    assert md.adapt(0.5, 0.33) == pytest.approx(0.5)
    assert md.adapt(0.5, 2.0) == pytest.approx(0.5 * math.exp(0.1 * (1 - 0.33)))   # al capped at 1
    assert md.adapt(0.5, 0.0, fast=True) == pytest.approx(0.5 * math.exp(-0.33))
    assert md.adapt(0.99, 1.0, fast=True) == 1.0                                      # capped at 1


def test_masks():
    rng = np.random.default_rng(3)
    st = md.SliceTime(10, 30, sim_frac=0.25)
    m = np.stack([md.slicetime_mask(40, 50, st, rng) for _ in range(40)])
    assert not m[:, :10].any() and not m[:, 30:].any()
    assert abs(m[:, 10:30].mean() - 0.25) < 0.01
    tree = [md.TreeNode(list(range(50)), [1, 2]), md.TreeNode([3, 9, 27]), md.TreeNode([])]
    ft = md.FixedTree(1, sim_frac=0.6)
    m = np.stack([md.fixedtree_mask(40, 50, ft, tree, rng) for _ in range(40)])
    other = np.setdiff1d(np.arange(50), [3, 9, 27])
    assert not m[:, :, other].any() and abs(m[:, :, [3, 9, 27]].mean() - 0.6) < 0.03
    full = md.slicetime_mask(40, 50, md.SliceTime(0, 40, 1.0), rng)
    assert full.all()
