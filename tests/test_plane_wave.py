"""The plane-wave M2L (docs/CONVENTIONS.md:1305-1434): an approximate operator beside the
exact M2L, checked as the reference checks its own (against the direct M2L,
fmm-tables/src/plane_wave tests): the quadrature reproduces 1/|u| on its region, one pair's
translation (host restatement of the device steps) equals the dense M2L's on multipoles of
charges to the quadrature's tolerance, and the GPU FMM with plane-wave M2L equals the
oracle FMM (exact M2L) to that tolerance."""
import numpy as np
import pytest
from scipy.special import j0

import fmm_amd
from oracle import oracle as O


@pytest.mark.parametrize("p,eps", [(6, 1e-5), (8, 1e-6), (12, 1e-8)])
def test_quadrature_reproduces_inverse_distance(p, eps):
    lam, w, na, H = fmm_amd.plane_wave_quadrature(p, eps)
    assert H == int(np.sum(na)) // 2 and np.all(na % 2 == 0)
    rng = np.random.default_rng(1)
    z = rng.uniform(1.0, 4.0, 4000)
    rho = np.sqrt(32.0) * np.sqrt(rng.uniform(0.0, 1.0, 4000))
    approx = (w[None, :] * np.exp(-lam[None, :] * z[:, None]) * j0(lam[None, :] * rho[:, None])).sum(1)
    rel = np.abs(approx * np.sqrt(z * z + rho * rho) - 1.0)
    assert rel.max() < eps, rel.max()
    # the trapezoid rule in alpha: (1/N) sum_j e^{i lam rho cos(a_j - t)} = J0(lam rho) + aliases
    for k in range(0, len(lam), 7):
        a = 2 * np.pi * np.arange(na[k]) / na[k]
        for r in (1.0, 3.0, np.sqrt(18.0)):
            tr = np.mean(np.exp(1j * lam[k] * r * np.cos(a - 0.3)))
            assert abs(tr - j0(lam[k] * r)) < 1e-3 * eps + 1e-15


@pytest.mark.parametrize("p,eps", [(8, 1e-6), (12, 1e-8)])
def test_pair_translation_matches_dense_m2l(p, eps):
    """Multipoles of charges in the source box translated by the plane waves and by the
    dense table; their locals evaluated in the target box differ by the quadrature error,
    far below the truncation both share."""
    rng = np.random.default_rng(3)
    _, _, m2l = fmm_amd.tables(p)
    worst = 0.0
    for t in range(0, 316, 9):
        d = np.array(O.offset_of_index(t))
        src = rng.uniform(-1, 1, (30, 3))
        q = rng.normal(size=30)
        M = O.p2m(p, np.zeros(3), 1.0, src, q)
        Ld = m2l[t] @ M
        Lp = fmm_amd.plane_wave_m2l(p, eps, t, M)
        tg = rng.uniform(-1, 1, (20, 3))
        pd, gd = O.l2p(p, np.zeros(3), 1.0, Ld, tg)
        pp, gp = O.l2p(p, np.zeros(3), 1.0, Lp, tg)
        worst = max(worst, np.linalg.norm(pp - pd) / np.linalg.norm(pd), np.linalg.norm(gp - gd) / np.linalg.norm(gd))
    assert worst < 20 * eps, worst


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,p,leaf", [("cube", 8000, 8, 40), ("plummer", 8000, 10, 32)])
def test_plane_wave_fmm_matches_oracle(kind, n, p, leaf):
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    print(f"case plane-wave fmm {kind} n={n} p={p}", flush=True)
    rng = O.Rng(11)
    pts = rng.cube(n) if kind == "cube" else rng.plummer(n, (0.5, 0.5, 0.5), 0.05)
    q = O.Rng(5).normals(n)
    pw = fmm_amd.FmmBuilder(p).max_points_per_leaf(leaf).m2l("plane_wave").build(pts)
    dense = fmm_amd.FmmBuilder(p).max_points_per_leaf(leaf).build(pts)
    a, b = pw.evaluate(q), dense.evaluate(q)
    ref = O.Fmm(pts, p=p, leaf_points=leaf)
    ophi, ograd = ref.evaluate(q)
    eps = 10 ** (-(0.5 * p + 2))
    e_pw = np.linalg.norm(a.potential - ophi) / np.linalg.norm(ophi)
    e_dense = np.linalg.norm(b.potential - ophi) / np.linalg.norm(ophi)
    eg = np.linalg.norm(a.gradient - ograd) / np.linalg.norm(ograd)
    print(f"  plane wave vs oracle {e_pw:.2e} (grad {eg:.2e}); dense vs oracle {e_dense:.2e}", flush=True)
    assert e_dense < 1e-12
    assert e_pw < 10 * eps and eg < 20 * eps


def test_plane_wave_is_f64_only():
    pts = np.random.default_rng(0).uniform(0, 1, (500, 3))
    with pytest.raises(fmm_amd.FmmError):
        fmm_amd.FmmBuilder(6).precision("mixed").m2l("plane_wave").build(pts)
