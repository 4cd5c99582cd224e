"""The plan's per-point stages on the GPU (fmmh_plan_build_device): the same plan bit for
bit as the host build (octree/src/octree/implementation.rs:682-716,1014): box keys of
every level, leaves, every list with its offset indices, the points' leaves, order and
leaf-scaled coordinates, and an operator over it evaluates the same bits."""
import time

import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.nlevels == b.nlevels and a.nleaves == b.nleaves
    assert np.array_equal(a.domain()[0], b.domain()[0]) and a.domain()[1] == b.domain()[1]
    assert np.array_equal(a.leaf_keys(), b.leaf_keys())
    for l in range(a.nlevels):
        assert np.array_equal(a.box_keys(l), b.box_keys(l))
        for kind in ("near", "v", "w", "x"):
            for x, y in zip(a.list(kind, l), b.list(kind, l)):
                assert np.array_equal(x, y), (kind, l)
    for t in (False, True):
        assert np.array_equal(a.point_leaves(t), b.point_leaves(t))
        assert np.array_equal(a.point_starts(t), b.point_starts(t))
        ua, oa = a.leaf_coordinates(t)
        ub, ob = b.leaf_coordinates(t)
        assert np.array_equal(oa, ob) and np.array_equal(ua, ub)


@pytest.mark.parametrize("kind,n,leaf,near", [("cube", 30000, 32, None), ("mixed", 40000, 24, (84.5, 84.5)),
                                              ("separate", 20000, 16, None)])
def test_device_plan_equals_host_plan(kind, n, leaf, near):
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    print(f"case device plan {kind} n={n} leaf={leaf}", flush=True)
    rng = O.Rng(8)
    if kind == "cube":
        src, tgt = rng.cube(n), None
    elif kind == "mixed":
        src, tgt = np.vstack([rng.plummer(n // 2, (0.2, 0.7, 0.4), 0.02), rng.cube(n - n // 2)]), None
    else:
        src, tgt = rng.cube(n), np.vstack([rng.plummer(n // 4, (0.5, 0.5, 0.5), 0.1), rng.cube(n // 4)])
    host = fmm_amd.Plan(src, tgt, leaf_points=leaf, near_rule=near)
    dev = fmm_amd.Plan(src, tgt, leaf_points=leaf, near_rule=near, device=0)
    _same(host, dev)
    if kind == "mixed":
        q = O.Rng(3).normals(n)
        a = fmm_amd.Fmm(host, fmm_amd.Operator(host, 6)).evaluate(q)
        b = fmm_amd.Fmm(dev, fmm_amd.Operator(dev, 6)).evaluate(q)
        assert np.array_equal(a.potential, b.potential) and np.array_equal(a.gradient, b.gradient)


def test_device_plan_at_the_headline_size():
    """4M uniform points, <= 128 per leaf, the device near rule: the same plan, and both
    build times."""
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    from fmm_amd import points as P

    pts, _, _ = P.bench_cube(4_000_000, 16)
    near = (84.5, 84.5)
    fmm_amd.Plan(pts[:1000], leaf_points=128, device=0)  # the first launches of the process
    t0 = time.perf_counter()
    host = fmm_amd.Plan(pts, leaf_points=128, near_rule=near)
    t1 = time.perf_counter()
    dev = fmm_amd.Plan(pts, leaf_points=128, near_rule=near, device=0)
    t2 = time.perf_counter()
    print(f"  plan build: host {t1 - t0:.2f} s, device stages {t2 - t1:.2f} s", flush=True)
    _same(host, dev)
