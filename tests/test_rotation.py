"""Rotation M2L (fmm-kernels/src/rotation.rs:441; CONVENTIONS §3.11 "Rotation of
coefficients", "Coaxial translations"): rotate the multipole so the offset lies on +z,
the coaxial shift, rotate the local back.

CPU: the composition G(Q^T) A(|d|) F(Q) of the library's rotation tables equals the dense
M2L table of the offset (the same operator) for every offset.  GPU: the rotation M2L of
an evaluation equals the dense one (f64 to rounding, f32 to its precision), and the whole
FMM with it matches the oracle.
"""
import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib
from oracle import oracle as O


@pytest.mark.parametrize("p", [3, 8, 12])
def test_rotation_tables_compose_to_dense(p):
    L = _lib.lib()
    nc = (p + 1) ** 2
    dense, m2m, l2l = np.zeros(316 * nc * nc), np.zeros(8 * nc * nc), np.zeros(8 * nc * nc)
    _lib.check(L.fmmh_tables(p, _lib.ptr(m2m), _lib.ptr(l2l), _lib.ptr(dense)))
    offsets = range(316) if p <= 8 else range(0, 316, 7)
    worst = 0.0
    for t in offsets:
        T = np.zeros(nc * nc)
        _lib.check(L.fmmh_m2l_rotation_table(p, t, _lib.ptr(T)))
        D = dense[t * nc * nc:(t + 1) * nc * nc]
        worst = max(worst, float(np.abs(T - D).max() / np.abs(D).max()))
    assert worst < 2e-12, worst


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bar", [("f64", 1e-12), ("mixed", 2e-6)])
def test_rotation_m2l_equals_dense(_gpu, precision, bar):
    rng = O.Rng(71)
    pts = np.vstack([rng.cube(15000), rng.plummer(8000, (0.4, 0.5, 0.6), 0.05)])
    q = rng.normals(len(pts))
    dense = fmm_amd.FmmBuilder(10).max_points_per_leaf(32).build(pts)
    rot = fmm_amd.FmmBuilder(10).precision(precision).m2l("rotation").max_points_per_leaf(32).build(pts)
    L = dense.plan.nlevels - 1
    locs = []
    for f in (dense, rot):
        op = f.operator
        op.load_charges(q)
        op.reset()
        for level in range(L, -1, -1):
            op.call("p2m", level)
        for level in range(L - 1, -1, -1):
            op.call("m2m", level)
        for level in range(2, L + 1):
            op.call("m2l", level)
        locs.append([op.expansions(True, level) for level in range(L + 1)])
    for level in range(2, L + 1):
        e = _rel(locs[1][level], locs[0][level])
        print(f"rotation m2l {precision} level {level}: rel {e:.3e}", flush=True)
        assert e < bar, level


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bar", [("f64", 1e-12), ("mixed", 1e-6)])
def test_rotation_fmm_matches_oracle(_gpu, precision, bar):
    # the whole FMM with the rotation M2L (f32 rotation blocks in the mixed precision)
    # against the oracle FMM (dense f64 M2L) on the same tree, with the near rule on
    rng = O.Rng(72)
    pts = np.vstack([rng.sphere(6000, (0.5, 0.5, 0.5), 0.45), rng.plummer(4000, (0.5, 0.5, 0.5), 0.1)])
    q = rng.charges(len(pts))
    near = (0.5 * 81,) * 2
    f = fmm_amd.FmmBuilder(8).precision(precision).m2l("rotation").max_points_per_leaf(32).near_rule(near).build(pts)
    out = f.evaluate(q)
    ref = O.Fmm(pts, p=8, leaf_points=32, near_rule=near)
    rphi, rgrad = ref.evaluate(q)
    e1, e2 = _rel(out.potential, rphi), _rel(out.gradient, rgrad)
    print(f"rotation fmm {precision} vs oracle: phi {e1:.3e} grad {e2:.3e}", flush=True)
    assert e1 < bar and e2 < bar


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bar", [("f64", 1e-12), ("mixed", 2e-6)])
def test_rotation_in_several_scratch_chunks_equals_dense(_gpu, precision, bar):
    # enough V pairs (1M points, p = 10: ~3M pairs) that the rotation M2L runs in several
    # chunks of scratch rows (512 MB each), each followed by its ordered reduction; the
    # result against the dense M2L on the same plan (itself pinned to the oracle)
    rng = O.Rng(95)
    pts = rng.cube(1_000_000)
    q = rng.normals(len(pts))
    plan = fmm_amd.Plan(pts, leaf_points=64)
    pairs = sum(len(plan.list("v", l)[1]) for l in range(plan.nlevels))
    per_chunk = (512 << 20) // (121 * (8 if precision == "f64" else 4))
    assert pairs > 2 * per_chunk
    rot = fmm_amd.Fmm(plan, fmm_amd.Operator(plan, 10, precision, True, 0, "rotation")).evaluate(q)
    dense = fmm_amd.Fmm(plan, fmm_amd.Operator(plan, 10, "f64", True, 0, "dense")).evaluate(q)
    e1 = O.error_norms(rot.potential, dense.potential)[0]
    e2 = O.error_norms(rot.gradient, dense.gradient, True)[0]
    print(f"rotation {precision} in {pairs / per_chunk:.1f} chunks vs dense f64: {e1:.2e} {e2:.2e}", flush=True)
    assert e1 < bar and e2 < bar
