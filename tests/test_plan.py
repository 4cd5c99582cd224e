"""CPU parity of the product's host side (tree, lists, leaf geometry, tables) with the
oracle, bit for bit where the quantities are integers or exact, and the C ABI surface.

No GPU is used: fmmh_plan_build and fmmh_tables are host functions of libfmm_amd.so.
"""
import json
import os
import re

import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "fmm_amd.h")).read()
    declared = set(re.findall(r"^[a-z_0-9 ]+\*?\s*(fmmh_[a-z0-9_]+)\s*\(", header, re.M))
    assert declared, "no declarations found"
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.fmmh_abi_version() == 6


def _problems():
    rng = O.Rng(101)
    yield "cube", rng.cube(3000), 20
    yield "plummer", rng.plummer(4000, (0.2, -0.1, 0.3), 0.05), 16
    yield "sphere", rng.sphere(4000, (0, 0, 0), 1.0), 12
    yield "clusters", rng.gaussian_clusters(3000, [[0, 0, 0], [1, 0.5, 0], [0.2, 0.9, 0.7]], 0.02), 10
    pts = rng.cube(2000)
    yield "duplicates", np.vstack([pts, pts[:500]]), 8


@pytest.mark.parametrize("name,pts,leaf", list(_problems()), ids=[p[0] for p in _problems()])
def test_plan_matches_oracle_bit_for_bit(name, pts, leaf):
    orc = O.Fmm(pts, p=1, leaf_points=leaf)
    plan = fmm_amd.Plan(pts, leaf_points=leaf)
    box_o, side_o = orc.domain()
    box_p, side_p = plan.domain()
    assert np.array_equal(box_o, box_p) and side_o == side_p
    assert plan.nlevels == orc.nlevels
    assert np.array_equal(plan.leaf_keys(), orc.leaf_keys())
    for level in range(plan.nlevels):
        assert np.array_equal(plan.box_keys(level), orc.box_keys(level)), level
        for kind in ("near", "w", "x"):
            po, pe = plan.list(kind, level)
            oo, oe = orc.list(kind, level)
            assert np.array_equal(po.astype(np.int64), oo), (kind, level)
            assert np.array_equal(pe.astype(np.int64), oe), (kind, level)
        po, pe, pa = plan.list("v", level)
        oo, oe, oa = orc.list("v", level)
        assert np.array_equal(po.astype(np.int64), oo)
        assert np.array_equal(pe.astype(np.int64), oe)
        assert np.array_equal(pa.astype(np.int64), oa)
    assert np.array_equal(plan.point_leaves().astype(np.int64), orc.point_leaves())
    up, order_p = plan.leaf_coordinates()
    uo, order_o = orc.leaf_coordinates()
    assert np.array_equal(order_p.astype(np.int64), order_o)
    assert np.array_equal(up, uo)  # same f64 operations, same bits


@pytest.mark.parametrize("rule", [(float("inf"), float("inf")), (10.0, 7.0), (0.0, 30.0)])
def test_near_rule_matches_oracle(rule):
    rng = O.Rng(77)
    pts = np.vstack([rng.plummer(3000, (0.5, 0.5, 0.5), 0.05), rng.cube(1500)])
    orc = O.Fmm(pts, p=1, leaf_points=12, near_rule=rule)
    plan = fmm_amd.Plan(pts, leaf_points=12, near_rule=rule)
    ref = fmm_amd.Plan(pts, leaf_points=12)
    moved = 0
    for level in range(plan.nlevels):
        for kind in ("near", "w", "x"):
            po, pe = plan.list(kind, level)
            oo, oe = orc.list(kind, level)
            assert np.array_equal(po.astype(np.int64), oo), (kind, level)
            assert np.array_equal(pe.astype(np.int64), oe), (kind, level)
        moved += len(plan.list("near", level)[1]) - len(ref.list("near", level)[1])
    if rule[0] > 0:
        assert moved > 0


def test_separate_targets_match_oracle():
    rng = O.Rng(7)
    src = rng.plummer(2500, (0.5, 0.5, 0.5), 0.08)
    tgt = rng.cube(1500)
    orc = O.Fmm(src, tgt, p=1, leaf_points=15)
    plan = fmm_amd.Plan(src, tgt, leaf_points=15)
    assert np.array_equal(plan.leaf_keys(), orc.leaf_keys())
    assert np.array_equal(plan.point_leaves(True).astype(np.int64), orc.point_leaves(True))
    ut, _ = plan.leaf_coordinates(True)
    uo, _ = orc.leaf_coordinates(True)
    assert np.array_equal(ut, uo)


def test_max_level_caps_the_tree():
    rng = O.Rng(9)
    pts = rng.plummer(3000, (0, 0, 0), 0.001)
    plan = fmm_amd.Plan(pts, leaf_points=4, max_level=5)
    orc = O.Fmm(pts, p=1, leaf_points=4, max_level=5)
    assert plan.nlevels == 6
    assert np.array_equal(plan.leaf_keys(), orc.leaf_keys())


def test_leaf_coordinates_lie_in_their_leaves():
    rng = O.Rng(3)
    pts = rng.cube(5000, (10.0, -4.0, 1e3), 2.5)
    plan = fmm_amd.Plan(pts, leaf_points=30)
    u, _ = plan.leaf_coordinates()
    assert np.max(np.abs(u)) <= 1.0 + 1e-9


@pytest.mark.parametrize("p", [2, 5, 8])
def test_tables_match_oracle(p):
    a, b, c = fmm_amd.tables(p)
    oa, ob, oc = O.tables(p)
    for mine, ref in ((a, oa), (b, ob), (c, oc)):
        scale = np.max(np.abs(ref), axis=(1, 2), keepdims=True)
        assert np.max(np.abs(mine - ref) / scale) < 1e-14


def test_tables_symmetry_at_high_degree():
    # M2L(-d) = diag((-1)^j) M2L(d) diag((-1)^n) (CONVENTIONS §3.12), at p = 13
    p = 13
    _, _, m2l = fmm_amd.tables(p)
    n = (p + 1) ** 2
    deg = np.floor(np.sqrt(np.arange(n))).astype(int)
    sgn = (-1.0) ** deg
    for d in ((2, 0, -3), (3, 3, 3), (-1, 2, 3)):
        a = m2l[O.offset_index(d)]
        b = m2l[O.offset_index(tuple(-v for v in d))]
        assert np.allclose(b, sgn[:, None] * a * sgn[None, :], rtol=1e-13, atol=0)


def test_uniform_tree_sizes_match_published():
    pub = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))
    pts, _, _ = O.bench_problem(1000000, 0)
    plan = fmm_amd.Plan(pts, leaf_points=64)
    row = [t for t in pub["trees"] if t["n"] == 1000000][0]
    assert plan.nleaves == row["leaves"]


@pytest.mark.slow
def test_ten_million_tree_matches_published():
    # fmm-bench/results/phase4s-gh200.md:64-66: 262,214 leaves on 8 levels at N = 10^7
    pub = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))
    row = [t for t in pub["trees"] if t["n"] == 10000000][0]
    pts, _, _ = O.bench_problem(10000000, 0)
    plan = fmm_amd.Plan(pts, leaf_points=64)
    assert plan.nleaves == row["leaves"]
    assert plan.nlevels == row["levels"]


def test_errors_are_reported():
    with pytest.raises(fmm_amd.FmmError) as e:
        fmm_amd.Plan(np.zeros((5, 3)))
    assert e.value.status == "InvalidArgument"
    with pytest.raises(fmm_amd.FmmError):
        fmm_amd.Plan(np.array([[0.0, 0, 0], [np.nan, 0, 0]]))
