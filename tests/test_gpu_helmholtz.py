"""GPU parity of the Helmholtz path (kernel e^{ikr}/(4 pi r), complex charges) through the
C ABI: the compensated direct sum against the oracle and the reference's FMM3D fixture;
the plain f64 FMM against the oracle Helmholtz FMM on the same tree (1e-12 relative L2) on
a uniform and an adaptive tree (W and X lists), with the dense M2L and with the per-pair
rotation M2L (the reference's RCR); the mixed precision (f32 P2P, bf16x3 M2L) against the
direct sum at the degree rule's p; real, complex and imaginary (Yukawa) k."""
import os
import sys

import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def _charges(n, seed=3):
    rng = O.Rng(seed)
    return rng.normals(n) + 1j * rng.normals(n)


def test_direct_sum_matches_fixture_and_oracle(golden_dir):
    _log("case helmholtz direct sum vs fmm3d fixture")
    z = np.load(os.path.join(golden_dir, "helmholtz_fmm3d.npz"))
    src, tgt = z["sources"], z["targets"]
    pts = np.vstack([src, tgt])
    for c in range(len(z["k"])):
        q = z["charges_real"] if z["charge_kind"][c] == 0 else z["charges_complex"]
        k = z["k"][c]
        phi, grad = fmm_amd.helmholtz_direct_sum(k, src, q, pts)
        assert _rel(phi / (4 * np.pi), z["potential"][c]) < 1e-13
        assert _rel(grad / (4 * np.pi), z["gradient"][c]) < 1e-13
        op, og = O.helmholtz_direct_sum(k, src, q, pts)
        assert _rel(phi, op) < 1e-14 and _rel(grad, og) < 1e-14


CASES = [
    ("uniform", 3000, 10, 40, 25.0),
    ("uniform", 3000, 8, 40, 5.0 + 0.5j),
    ("clustered", 3000, 10, 32, 12.0),
    ("clustered", 2000, 8, 32, 3.0j),
]


def _points(kind, n, seed=17):
    rng = O.Rng(seed)
    if kind == "uniform":
        return rng.cube(n)
    return np.vstack([rng.plummer(n // 2, (0.4, 0.5, 0.5), 0.04), rng.cube(n - n // 2)])


@pytest.mark.parametrize("kind,n,p,leaf,k", CASES)
def test_plain_f64_matches_oracle_fmm(kind, n, p, leaf, k):
    _log(f"case helmholtz plain f64 {kind} n={n} p={p} leaf={leaf} k={k}")
    pts = _points(kind, n)
    q = _charges(n)
    fmm = fmm_amd.HelmholtzFmmBuilder(k, degree=p).max_points_per_leaf(leaf).build(pts)
    if kind == "clustered":
        L = fmm.plan.nlevels
        assert any(len(fmm.plan.list("w", l)[1]) for l in range(L))
        assert any(len(fmm.plan.list("x", l)[1]) for l in range(L))
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=p, leaf_points=leaf)
    assert ref.nlevels == fmm.plan.nlevels
    ophi, ograd = ref.evaluate_helmholtz(k, q)
    ep, eg = _rel(out.potential, ophi), _rel(out.gradient, ograd)
    _log(f"  vs oracle: phi {ep:.2e} grad {eg:.2e}")
    assert ep < 1e-12 and eg < 1e-12
    # and the oracle itself against the direct sum at a sample (the FMM's truncation)
    s = np.arange(0, n, 37)
    dp, dg = fmm_amd.helmholtz_direct_sum(k, pts, q, pts[s])
    _log(f"  vs direct: phi {_rel(out.potential[s], dp / (4 * np.pi)):.2e}")
    assert _rel(out.potential[s], dp / (4 * np.pi)) < 1e-3


@pytest.mark.parametrize("kind,n,p,leaf,k", CASES)
def test_rotation_m2l_matches_oracle_fmm(kind, n, p, leaf, k):
    """The per-pair rotation M2L (rotation, coaxial shift, rotation back in f64) against the
    oracle Helmholtz FMM (1e-12) and the dense path on the same plan."""
    _log(f"case helmholtz rotation m2l {kind} n={n} p={p} leaf={leaf} k={k}")
    pts = _points(kind, n)
    q = _charges(n)
    rot = fmm_amd.HelmholtzFmmBuilder(k, degree=p).max_points_per_leaf(leaf).m2l("rotation").build(pts)
    dense = fmm_amd.HelmholtzFmm(rot.plan, fmm_amd.HelmholtzOperator(rot.plan, p, k))
    out, d = rot.evaluate(q), dense.evaluate(q)
    ref = O.Fmm(pts, p=p, leaf_points=leaf)
    ophi, ograd = ref.evaluate_helmholtz(k, q)
    ep, eg = _rel(out.potential, ophi), _rel(out.gradient, ograd)
    _log(f"  vs oracle: phi {ep:.2e} grad {eg:.2e}; vs dense {_rel(out.potential, d.potential):.2e}")
    assert ep < 1e-12 and eg < 1e-12
    assert _rel(out.potential, d.potential) < 1e-12 and _rel(out.gradient, d.gradient) < 1e-12
    # the translation tables: O(p^3) per distance instead of O(p^4) per offset and level
    assert rot.operator.stats()["device_bytes"] < dense.operator.stats()["device_bytes"]


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_dense_budget_mixes_dense_and_rotation_levels(precision):
    """HelmholtzM2l::Dense under a byte budget (DenseM2l::select): dense tables on the finest
    levels that fit, per-pair rotation on the coarser ones; the mix against every level
    dense and against the oracle."""
    n, p, leaf, k = 6000, 10, 32, 14.0 + 0.2j
    _log(f"case helmholtz dense budget n={n} p={p} {precision}")
    pts = _points("uniform", n)
    q = _charges(n)
    level = 316 * 2 * (p + 1) ** 4 * 8
    b = fmm_amd.HelmholtzFmmBuilder(k, degree=p).max_points_per_leaf(leaf).precision(precision)
    mix = b.dense_budget(level).build(pts)  # the finest level dense, the others rotation
    assert mix.plan.nlevels >= 4
    full = fmm_amd.HelmholtzFmm(mix.plan, fmm_amd.HelmholtzOperator(mix.plan, p, k, precision, dense_budget=-1))
    a, f = mix.evaluate(q), full.evaluate(q)
    bar = 1e-12 if precision == "f64" else 1e-6
    _log(f"  mix vs all dense: phi {_rel(a.potential, f.potential):.2e} grad {_rel(a.gradient, f.gradient):.2e}")
    assert _rel(a.potential, f.potential) < bar and _rel(a.gradient, f.gradient) < bar
    if precision == "f64":
        ophi, ograd = O.Fmm(pts, p=p, leaf_points=leaf).evaluate_helmholtz(k, q)
        assert _rel(a.potential, ophi) < 1e-12 and _rel(a.gradient, ograd) < 1e-12


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_rotation_m2l_at_the_degree_rule(precision):
    """The rotation M2L at the degree rule for 1e-6 (p = 15 at k = 25 here, the 256-point
    leaves of the reference's default): within the tolerance of the direct sum, and equal
    to the dense path's plain f64 result to 1e-10."""
    n, k = 20000, 25.0
    _log(f"case helmholtz rotation m2l at the degree rule n={n} {precision}")
    pts = _points("uniform", n)
    q = _charges(n)
    b = fmm_amd.HelmholtzFmmBuilder(k, tolerance=1e-6).near_wx("device").precision(precision)
    rot = b.m2l("rotation").build(pts)
    s = np.arange(0, n, 97)
    dp, dg = fmm_amd.helmholtz_direct_sum(k, pts, q, pts[s])
    out = rot.evaluate(q)
    ep, eg = _rel(out.potential[s], dp / (4 * np.pi)), _rel(out.gradient[s], dg / (4 * np.pi))
    _log(f"  p={rot.p}: vs direct phi {ep:.2e} grad {eg:.2e}")
    assert ep < 1e-6 and eg < 1e-6
    if precision == "f64":
        dense = fmm_amd.HelmholtzFmm(rot.plan, fmm_amd.HelmholtzOperator(rot.plan, rot.p, k))
        d = dense.evaluate(q)
        assert _rel(out.potential, d.potential) < 1e-10 and _rel(out.gradient, d.gradient) < 1e-10


def test_near_rule_matches_oracle_fmm():
    """W and X pairs converted to P2P (WxNear::Crossovers) on an adaptive tree: the same
    plan in the oracle (its near rule) and on the GPU agree to 1e-12."""
    n, p, leaf, k = 3000, 8, 32, 9.0 + 0.3j
    _log(f"case helmholtz near rule n={n} p={p}")
    pts = _points("clustered", n)
    q = _charges(n)
    rule = (20.0, 20.0)
    fmm = fmm_amd.HelmholtzFmmBuilder(k, degree=p).max_points_per_leaf(leaf).near_wx(rule).build(pts)
    off = fmm_amd.Plan(pts, leaf_points=leaf)
    L = off.nlevels
    assert sum(len(fmm.plan.list("w", l)[1]) for l in range(L)) < sum(len(off.list("w", l)[1]) for l in range(L))
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=p, leaf_points=leaf, near_rule=rule)
    ophi, ograd = ref.evaluate_helmholtz(k, q)
    assert _rel(out.potential, ophi) < 1e-12 and _rel(out.gradient, ograd) < 1e-12


def test_evaluations_repeat_bit_for_bit():
    _log("case helmholtz repeat")
    pts = _points("uniform", 2000)
    q = _charges(2000)
    fmm = fmm_amd.HelmholtzFmmBuilder(7.0, degree=8).max_points_per_leaf(32).build(pts)
    a = fmm.evaluate(q)
    b = fmm.evaluate(q)
    assert np.array_equal(a.potential, b.potential) and np.array_equal(a.gradient, b.gradient)
    # linearity in the charges
    c = fmm.evaluate(2.0 * q)
    assert np.allclose(c.potential, 2.0 * a.potential, rtol=1e-13, atol=0)


@pytest.mark.parametrize("k", [25.0, 10.0 + 1.0j])
def test_mixed_precision_within_tolerance(k):
    """The mixed path at the degree rule for 1e-6 (uniform, the 256-point leaves of the
    reference's default): relative L2 of phi and grad phi against the direct sum."""
    n = 20000
    _log(f"case helmholtz mixed n={n} k={k}")
    pts = _points("uniform", n)
    q = _charges(n)
    plain = fmm_amd.HelmholtzFmmBuilder(k, tolerance=1e-6).near_wx("device").build(pts)
    mixed = fmm_amd.HelmholtzFmmBuilder(k, tolerance=1e-6).near_wx("device").precision("mixed").build(pts)
    assert mixed.p == plain.p
    s = np.arange(0, n, 97)
    dp, dg = fmm_amd.helmholtz_direct_sum(k, pts, q, pts[s])
    dp, dg = dp / (4 * np.pi), dg / (4 * np.pi)
    for name, f, bar in (("plain", plain, 1e-6), ("mixed", mixed, 2e-6)):
        out = f.evaluate(q)
        ep, eg = _rel(out.potential[s], dp), _rel(out.gradient[s], dg)
        _log(f"  {name} p={f.p}: phi {ep:.2e} grad {eg:.2e}")
        assert ep < bar and eg < bar


@pytest.mark.parametrize("kind,n,p,leaf,k", [("uniform", 8000, 10, 32, 25.0), ("clustered", 8000, 12, 32, 14.0 + 0.2j),
                                              ("uniform", 8000, 10, 32, 12.0 + 8.0j)])
def test_plane_wave_m2l_matches_oracle_fmm(kind, n, p, leaf, k):
    """The exponential M2L by merge and shift (HelmholtzM2l::PlaneWave: the levels the rule
    gives a quadrature, RCR on the others) against the oracle FMM (exact M2L) and the
    direct sum: within the rotation path's error of the direct sum (the reference finds
    iprec 2 within 1.05x of Rotation's) and within that error of the oracle.  k = 12 + 8i
    puts the coarse levels in FMM3D's real-t domains (Im kappa >= pi/2), the others on
    the contour."""
    _log(f"case helmholtz plane-wave m2l {kind} n={n} p={p} leaf={leaf} k={k}")
    pts = _points(kind, n)
    q = _charges(n)
    pw = fmm_amd.HelmholtzFmmBuilder(k, degree=p).max_points_per_leaf(leaf).m2l("plane_wave").build(pts)
    rot = fmm_amd.HelmholtzFmm(pw.plan, fmm_amd.HelmholtzOperator(pw.plan, p, k, m2l="rotation"))
    a, r = pw.evaluate(q), rot.evaluate(q)
    ophi, ograd = O.Fmm(pts, p=p, leaf_points=leaf).evaluate_helmholtz(k, q)
    s = np.arange(0, n, 29)
    dp, dg = fmm_amd.helmholtz_direct_sum(k, pts, q, pts[s])
    dp, dg = dp / (4 * np.pi), dg / (4 * np.pi)
    t_pw, t_rot = _rel(a.potential[s], dp), _rel(r.potential[s], dp)
    g_pw, g_rot = _rel(a.gradient[s], dg), _rel(r.gradient[s], dg)
    e_pw, eg_pw = _rel(a.potential, ophi), _rel(a.gradient, ograd)
    _log(f"  vs direct: phi {t_pw:.2e} (rotation {t_rot:.2e}) grad {g_pw:.2e} (rotation {g_rot:.2e}); "
         f"vs oracle {e_pw:.2e} / {eg_pw:.2e}")
    assert _rel(r.potential, ophi) < 1e-12
    assert t_pw <= 1.5 * t_rot and g_pw <= 1.5 * g_rot
    assert e_pw <= t_rot and eg_pw <= g_rot


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_plane_wave_m2l_at_the_degree_rule(precision):
    """The exponential M2L at the degree rule for 1e-6 (p = 15 at k = 25, iprec 2 on every
    level with a V list): within the tolerance of the direct sum."""
    n, k = 20000, 25.0
    _log(f"case helmholtz plane-wave m2l at the degree rule n={n} {precision}")
    pts = _points("uniform", n)
    q = _charges(n)
    b = fmm_amd.HelmholtzFmmBuilder(k, tolerance=1e-6).near_wx("device").precision(precision)
    pw = b.m2l("plane_wave").build(pts)
    s = np.arange(0, n, 97)
    dp, dg = fmm_amd.helmholtz_direct_sum(k, pts, q, pts[s])
    out = pw.evaluate(q)
    ep, eg = _rel(out.potential[s], dp / (4 * np.pi)), _rel(out.gradient[s], dg / (4 * np.pi))
    _log(f"  p={pw.p}: vs direct phi {ep:.2e} grad {eg:.2e}")
    assert ep < 1e-6 and eg < 1e-6
