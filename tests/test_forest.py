"""Batched independent clouds as one plan (a forest).

CPU: the forest is each cloud's own plan stacked level by level -- box keys, leaves,
lists (indices shifted by the clouds before it) and leaf-scaled coordinates equal the
single-cloud plans'.  GPU: one operator on the forest gives each cloud what its own
operator gives (f64 to rounding; the pair/tile choice of a level may differ), and the
adjoint works on the forest.
"""
import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O


def _clouds(seed=51):
    rng = O.Rng(seed)
    return [
        rng.gaussian_clusters(4000, [(0.5, 0.5, 0.5)], 0.1),
        rng.cube(700),
        rng.gaussian_clusters(9000, [(0.3, 0.3, 0.6), (0.7, 0.6, 0.4)], 0.05),
        rng.plummer(2500, (0.5, 0.5, 0.5), 0.05),
    ]


def test_forest_stacks_the_cloud_plans():
    clouds = _clouds()
    leaf = 24
    forest = fmm_amd.Plan.forest(clouds, leaf_points=leaf)
    single = [fmm_amd.Plan(c, leaf_points=leaf) for c in clouds]
    L = max(s.nlevels for s in single)
    assert forest.nlevels == L
    assert forest.nleaves == sum(s.nleaves for s in single)
    lk = forest.leaf_keys()
    lvl = (lk & np.uint64(0x7FFF)).astype(int)
    s_lk = [s.leaf_keys() for s in single]
    s_lvl = [(k & np.uint64(0x7FFF)).astype(int) for k in s_lk]
    # global leaf index of each cloud's leaves
    leaf_map = [np.zeros(s.nleaves, np.int64) for s in single]
    for l in range(L):
        at = int(np.searchsorted(lvl, l))
        for c, s in enumerate(single):
            idx = np.nonzero(s_lvl[c] == l)[0]
            leaf_map[c][idx] = at + np.arange(len(idx))
            at += len(idx)
    for c in range(len(single)):
        assert np.array_equal(lk[leaf_map[c]], s_lk[c])
    for l in range(L):
        keys = [s.box_keys(l) if l < s.nlevels else np.zeros(0, np.uint64) for s in single]
        assert np.array_equal(forest.box_keys(l), np.concatenate(keys))
        base = np.concatenate([[0], np.cumsum([len(k) for k in keys])])
        off, ent, aux = forest.list("v", l)
        ents, auxs, offs = [], [], [0]
        for c, s in enumerate(single):
            if l >= s.nlevels:
                continue
            o, e, a = s.list("v", l)
            ents.append(e.astype(np.int64) + base[c])
            auxs.append(a)
            offs.extend((o[1:].astype(np.int64) + offs[-1]).tolist())
        assert np.array_equal(off, np.array(offs, np.uint32))
        if ents:
            assert np.array_equal(ent, np.concatenate(ents).astype(np.uint32))
            assert np.array_equal(aux, np.concatenate(auxs))
        off, ent = forest.list("near", l)
        ents = [leaf_map[c][s.list("near", l)[1]] for c, s in enumerate(single) if l < s.nlevels]
        if ents:
            assert np.array_equal(ent, np.concatenate(ents).astype(np.uint32))
    # points: the leaf of every point and its leaf-scaled coordinates
    pl = forest.point_leaves()
    u, order = forest.leaf_coordinates()  # leaf order, with the caller index of each
    uc = np.empty_like(u)
    uc[order] = u
    first = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    for c, s in enumerate(single):
        assert np.array_equal(pl[first[c]:first[c + 1]], leaf_map[c][s.point_leaves()])
        su, so = s.leaf_coordinates()
        sc = np.empty_like(su)
        sc[so] = su
        assert np.array_equal(uc[first[c]:first[c + 1]], sc)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_forest_evaluates_each_cloud(precision):
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    clouds = _clouds(52)
    rng = O.Rng(53)
    qs = [rng.normals(len(c)) for c in clouds]
    b = fmm_amd.FmmBuilder(8).precision(precision).max_points_per_leaf(32)
    forest = b.build_batched(clouds)
    out = forest.evaluate(np.concatenate(qs))
    first = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    ref = fmm_amd.FmmBuilder(8).max_points_per_leaf(32)
    for c, (pts, q) in enumerate(zip(clouds, qs)):
        alone = ref.build(pts).evaluate(q)
        e1 = _rel(out.potential[first[c]:first[c + 1]], alone.potential)
        e2 = _rel(out.gradient[first[c]:first[c + 1]], alone.gradient)
        print(f"forest {precision} cloud {c} (n={len(pts)}): rel phi {e1:.3e} grad {e2:.3e}", flush=True)
        bar = 1e-13 if precision == "f64" else 1e-6
        assert e1 < bar and e2 < bar
    # the adjoint on the forest: each cloud's own adjoint
    g = 2.0 * out.potential / len(out.potential)
    gq, gz = forest.vjp(np.concatenate(qs), g, out.gradient)
    for c, (pts, q) in enumerate(zip(clouds, qs)):
        f1 = ref.build(pts)
        o1 = f1.evaluate(q)
        sl = slice(first[c], first[c + 1])
        rq, rz = f1.vjp(q, g[sl], o1.gradient)
        bar = 1e-12 if precision == "f64" else 2e-6
        assert _rel(gq[sl], rq) < bar and _rel(gz[sl], rz) < bar
