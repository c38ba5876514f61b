"""Slice-time and fixed-tree MBP proposals and the geographic tree behind
the fixed-tree ones.

Mbp::mbp_slicetime / Mbp::mbp_fixedtree (src/mbp.cc:656-716) keep a chain's
parameters and re-draw part of its trajectory.  In the cells of a simulation
mask (step x area) the transitions are drawn afresh, everywhere else they
are coupled to the current trajectory (the MBP of beep_mbp).  The mask of a
slice-time proposal marks each (step, area) of its time window
[sett_i, sett_f) with probability sim_frac; that of a fixed-tree proposal
marks every step of the areas of one tree node with probability sim_frac.
After the Metropolis-Hastings step the proposal's sim_frac adapts towards the
target acceptance rate, sim_frac *= exp(eta (min(al, 1) - 0.33)), capped at 1
(update, src/utils.cc:619-623; eta 1 for fast updates and 0.1 for slow ones,
src/consts.hh:137-142), and its trial and acceptance counts grow (ntr += 1,
nac += al).  The reference's default proposal list holds one slice-time
proposal over every step with sim_frac 1 (ParamProp::ParamProp,
src/param_prop.cc:36-41); its fixed-tree list is empty there.

Data::area_split (src/data_matrix.cc:299-437) builds the tree: the root
holds every area; a node of more than one area is split in two by swapping
pairs of areas between the halves while the split's fit improves, stopping
after 50 swaps in a row that do not improve it; the fit rewards mixing
weight within a half and penalises weight across (G[j][j] = M_jj pop_j,
G[j][i] = M_ji sqrt(pop_j pop_i) on the normalised geographic matrix).

The masks and swaps draw on a numpy Generator, not the reference's ran()
stream: the mask of a proposal is a function of its uniforms, and which
uniforms a run sees is not part of the algorithm.
"""
import math
from dataclasses import dataclass, field

import numpy as np

ETA_FAST, ETA_SLOW = 1.0, 0.1          # src/consts.hh:137
AC_RATE = 0.33                         # src/consts.hh:142, the target acceptance rate


@dataclass
class SliceTime:
    """A slice-time proposal: steps [sett_i, sett_f) (struct SliceTime)."""
    sett_i: int
    sett_f: int
    sim_frac: float = 1.0
    ntr: int = 0
    nac: float = 0.0


@dataclass
class FixedTree:
    """A fixed-tree proposal on tree node n (struct FixedTree)."""
    n: int
    sim_frac: float = 1.0
    ntr: int = 0
    nac: float = 0.0


@dataclass
class TreeNode:
    arearef: list
    child: list = field(default_factory=list)


def adapt(sim_frac, al, fast=False):
    """update(sim_frac, min(al, 1), eta) of Mbp::mbp_slicetime /
    mbp_fixedtree, capped at 1."""
    al = min(al, 1.0)
    v = sim_frac * math.exp((ETA_FAST if fast else ETA_SLOW) * (al - AC_RATE))
    return min(v, 1.0)


def normalised_mixing(spec):
    """The geographic matrix after Data::geo_normalise (src/data_matrix.cc:
    55-96): row c divided by sum_j M_cj pop_j (diagonal included); returns
    (diag [A], rowptr, to, val, area populations [A])."""
    A = spec["A"]
    pop = np.asarray(spec["pop"], dtype=np.float64).reshape(A, -1).sum(axis=1)
    M = spec.get("M")
    if M is None:
        if A != 1:
            raise ValueError("a geographic mixing matrix is needed when there is more than one area")
        return np.ones(1), np.zeros(2, dtype=np.int64), np.zeros(0, dtype=np.int64), np.zeros(0), pop
    diag = np.asarray(M["diag"], dtype=np.float64).copy()
    rowptr = np.asarray(M["rowptr"], dtype=np.int64)
    to = np.asarray(M["to"], dtype=np.int64)
    val = np.asarray(M["val"], dtype=np.float64).copy()
    for c in range(A):
        if diag[c] == 0:
            raise ValueError("the geographic mixing matrix cannot have zero on the diagonal")
        a, b = rowptr[c], rowptr[c + 1]
        s = diag[c] * pop[c] + float((val[a:b] * pop[to[a:b]]).sum())
        diag[c] /= s
        val[a:b] /= s
    return diag, rowptr, to, val, pop


def split_matrix(spec):
    """G of Data::area_split: G[j][j] = M_jj pop_j, G[j][i] = M_ji
    sqrt(pop_j pop_i) (dense, [A][A])."""
    diag, rowptr, to, val, pop = normalised_mixing(spec)
    A = len(diag)
    G = np.zeros((A, A))
    for j in range(A):
        G[j, j] = diag[j] * pop[j]
        for k in range(rowptr[j], rowptr[j + 1]):
            i = to[k]
            if i == j:
                raise ValueError("a geographic mixing entry on the diagonal of the off-diagonal list")
            G[j, i] = val[k] * math.sqrt(pop[j] * pop[i])
    return G


def split_fit(area_list, gr, G):
    """Data::get_split_fit: sum over ordered pairs i != j of G[a_j][a_i],
    added within a half and subtracted across."""
    a = np.asarray(area_list)
    sub = G[np.ix_(a, a)].T                          # sub[i][j] = G[a_j][a_i]
    same = np.equal.outer(gr, gr)
    np.fill_diagonal(sub, 0.0)
    return float(np.where(same, sub, -sub).sum())


def split_fit_dif(area_list, gr, G, isel):
    """Data::get_split_fit_dif: the change of the fit when area isel changes
    half; flips gr[isel]."""
    a = np.asarray(area_list)
    v = G[a[isel], a] + G[a, a[isel]]
    v[isel] = 0.0
    same = gr == gr[isel]
    d = float(np.where(same, -2.0 * v, 2.0 * v).sum())
    gr[isel] = not gr[isel]
    return d


def split_in_two(area_list, G, rng, max_fail=50):
    """Data::split_in_two: the first half of the list against the second,
    then random swaps of an area from each half, kept when the fit rises,
    until max_fail swaps in a row do not."""
    si = len(area_list)
    gr = np.arange(si) < si // 2
    fit = split_fit(area_list, gr, G)
    fail = 0
    while fail < max_fail:
        while True:
            i, j = int(rng.random() * si), int(rng.random() * si)
            if i != j and gr[i] != gr[j]:
                break
        new = fit + split_fit_dif(area_list, gr, G, i)
        new += split_fit_dif(area_list, gr, G, j)
        if new > fit:
            fail = 0
            fit = new
        else:
            fail += 1
            gr[i], gr[j] = gr[j], gr[i]
    ch1 = [area_list[i] for i in range(si) if gr[i]]
    ch2 = [area_list[i] for i in range(si) if not gr[i]]
    return ch1, ch2


def area_split(spec, rng):
    """Data::area_split: the tree of nodes (breadth first from the root of
    all areas), every node of two or more areas split in two."""
    G = split_matrix(spec)
    nodes = [TreeNode(list(range(spec["A"])))]
    n = 0
    while n < len(nodes):
        if len(nodes[n].arearef) > 1:
            c1, c2 = split_in_two(nodes[n].arearef, G, rng)
            nodes[n].child.append(len(nodes))
            nodes.append(TreeNode(c1))
            nodes[n].child.append(len(nodes))
            nodes.append(TreeNode(c2))
        n += 1
    return nodes


def slicetime_mask(S, A, st, rng):
    """The simulation mask of a slice-time proposal: [S][A] uint8."""
    m = np.zeros((S, A), dtype=np.uint8)
    w = rng.random((st.sett_f - st.sett_i, A)) < st.sim_frac
    m[st.sett_i:st.sett_f] = w
    return m


def fixedtree_mask(S, A, ft, tree, rng):
    """The simulation mask of a fixed-tree proposal on node ft.n: [S][A]
    uint8, the node's areas marked step by step with probability sim_frac."""
    m = np.zeros((S, A), dtype=np.uint8)
    areas = np.asarray(tree[ft.n].arearef)
    m[:, areas] = rng.random((S, len(areas))) < ft.sim_frac
    return m


def propose(ctx, init, prop, slot, proposals, key, rng, seed=1, stream=0, invT=1.0, tree=None, fast=False,
            adapt_frac=True):
    """One slice-time or fixed-tree proposal per chain, all at once on the
    device: chain i's current state is instance init[i] (parameters in slot
    slot[i], its EF from the last observation log-likelihood on it) and its
    proposal is written to instance prop[i] with proposals[i] (a SliceTime or
    FixedTree, adapted in place).  The MBPs with their masks, the proposals'
    observation log-likelihoods and the Metropolis-Hastings steps are each
    one call (Mbp::get_al at inverse temperature invT; the parameters do not
    change, so the prior ratio is 0).  Returns (al, accepted): accepted[i]
    means prop[i] now holds chain i's state."""
    n = len(init)
    S, A = ctx.S, ctx.A
    masks = np.zeros((n, S, A), dtype=np.uint8)
    for i, pp in enumerate(proposals):
        if isinstance(pp, SliceTime):
            if not 0 <= pp.sett_i < pp.sett_f <= S:
                raise ValueError("slice-time window out of range")
            masks[i] = slicetime_mask(S, A, pp, rng)
        elif isinstance(pp, FixedTree):
            if tree is None or not 0 <= pp.n < len(tree):
                raise ValueError("a fixed-tree proposal needs the area tree and a node in it")
            masks[i] = fixedtree_mask(S, A, pp, tree, rng)
        else:
            raise TypeError("proposals are SliceTime or FixedTree")
    ctx.mbp(init, prop, slot, key, simu_mask=masks, seed=seed, stream=stream)
    ctx.obs_loglik(prop)
    al, acc = ctx.accept(init, prop, 1, np.zeros(n), key, invT=invT, seed=seed, stream=stream)
    for i, pp in enumerate(proposals):
        a = min(float(al[i]), 1.0)
        if adapt_frac:
            pp.sim_frac = adapt(pp.sim_frac, a, fast)
        pp.ntr += 1
        pp.nac += a
    return al, acc
