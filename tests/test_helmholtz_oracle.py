"""Pins of the Helmholtz oracle (CPU): the basis against the addition theorem, the radial
functions against scipy, the coaxial recursion against the translated wave functions
evaluated pointwise, the rotation blocks against rotated harmonics, every translation chain
and the whole oracle FMM against f64 direct sums (the reference's checks of its Helmholtz
operators, fmm-ref/tests/helmholtz and docs/CONVENTIONS.md:1435-1700)."""
import numpy as np
import pytest
from scipy.special import eval_legendre, spherical_jn, spherical_yn

from oracle import oracle as O


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_orthonormal_harmonics_addition_theorem():
    rng = np.random.default_rng(1)
    p = 14
    for _ in range(3):
        x, y = rng.normal(size=(2, 3))
        x /= np.linalg.norm(x)
        y /= np.linalg.norm(y)
        sx, sy = O.h_solid(p, x), O.h_solid(p, y)
        for n in range(p + 1):
            s = float(np.dot(sx[n * n:(n + 1) ** 2], sy[n * n:(n + 1) ** 2]))
            assert abs(s - (2 * n + 1) / (4 * np.pi) * eval_legendre(n, x @ y)) < 1e-13 * (2 * n + 1)
    # Y_1: sqrt(3/4pi) (x, y, z) in slots (+1, -1, 0); no Condon-Shortley phase
    v = np.array([0.3, -0.5, 0.8])
    s = O.h_solid(1, v)
    c = np.sqrt(3 / (4 * np.pi))
    assert np.allclose(s[1:], c * np.array([v[1], v[2], v[0]]), rtol=0, atol=1e-15)


def test_solid_gradient_by_finite_differences():
    x = np.array([0.4, -0.7, 0.25])
    p, h = 8, 1e-6
    g = O.h_solid_grad(p, x)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd = (O.h_solid(p, x + e) - O.h_solid(p, x - e)) / (2 * h)
        assert np.max(np.abs(fd - g[k])) < 1e-8


@pytest.mark.parametrize("z", [1e-3, 0.3, 2.5, 7.0, 12.0, 30.0, 1 + 0.5j, 3.9 + 0.6j, 0.05j, 2j])
def test_radial_functions_match_scipy(z):
    P = 48
    j, h = O.h_radial(P, z)
    n = np.arange(P + 1)
    jr = spherical_jn(n, z)
    hr = jr + 1j * spherical_yn(n, z)
    assert np.max(np.abs(j - jr) / np.abs(jr)) < 1e-12
    assert np.max(np.abs(h - hr) / np.abs(hr)) < 1e-12


def _waves(kind, p, k, x):
    r = np.linalg.norm(x)
    n = np.arange(p + 1)
    f = spherical_jn(n, k * r) if kind == "R" else spherical_jn(n, k * r) + 1j * spherical_yn(n, k * r)
    y = O.h_solid(p, x / r)
    return np.repeat(f, 2 * n + 1) * y


@pytest.mark.parametrize("k", [0.7, 1.3 + 0.2j, 4.0])
@pytest.mark.parametrize("kind", ["SR", "RR", "SS"])
def test_coaxial_recursion_reproduces_translated_waves(k, kind):
    """F_n^m(r + t z) = sum_n' (E|F)^m_{n'n}(t) G_n'^m(r) pointwise (CONVENTIONS:1632-1645)."""
    rng = np.random.default_rng(2)
    P, t = 36, 3.0
    E = O.h_coaxial(P, kind == "SR", k, t)
    err = 0.0
    for _ in range(3):
        x = rng.normal(size=3)
        x *= (0.12 * t if kind != "SS" else 8 * t) / np.linalg.norm(x)
        lhs = _waves(kind[0], P, k, x + np.array([0, 0, t]))
        g = _waves(kind[1], P, k, x)
        for n in range(6):
            for m in range(-n, n + 1):
                am = abs(m)
                rhs = sum(E[am, q, n] * g[q * q + q + m] for q in range(am, P + 1))
                err = max(err, abs(rhs - lhs[n * n + n + m]) / abs(lhs[n * n + n + m]))
    assert err < (1e-9 if kind == "SS" else 1e-12), err


def test_coaxial_index_symmetry():
    E = O.h_coaxial(10, 1, 0.9 + 0.1j, 4.0)
    for m in range(11):
        for a in range(m, 11):
            for b in range(m, 11):
                assert E[m, a, b] == pytest.approx((-1) ** (a + b) * E[m, b, a], rel=1e-12, abs=0)


def test_rotation_blocks():
    rng = np.random.default_rng(3)
    A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    Q = A * np.sign(np.linalg.det(A))
    p = 10
    B = O.h_rotation_blocks(p, Q)
    for _ in range(3):
        x = rng.normal(size=3)
        a, b = O.h_solid(p, Q @ x), O.h_solid(p, x)
        for n in range(p + 1):
            s = slice(n * n, (n + 1) ** 2)
            assert np.max(np.abs(a[s] - B[n] @ b[s])) < 1e-13 * max(1.0, np.abs(a[s]).max())
            assert np.max(np.abs(B[n] @ B[n].T - np.eye(2 * n + 1))) < 1e-13


def _direct(k, src, q, tgt):
    return O.helmholtz_direct_sum(k, src, q, tgt)


@pytest.mark.parametrize("k", [0.4, 1.5, 1.0 + 0.3j, 0.05])
def test_translation_chains_match_direct_sums(k):
    """Every operator of the chain in level units against e^{ik r}/r summed directly:
    P2M-M2P, P2M-M2L-L2P, P2L-L2P, P2M-M2M-M2P, P2L-L2L-L2P (a dropped term, a wrong sign,
    scale or index in any of them fails by orders of magnitude)."""
    rng = np.random.default_rng(5)
    p = 14
    src = rng.uniform(-1, 1, (30, 3))
    q = rng.normal(size=30) + 1j * rng.normal(size=30)
    d = np.array([3, 1, -2])
    z3 = np.zeros(3)
    tol_phi, tol_grad = 1e-9, 5e-9
    # P2M -> M2P
    M = O.h_p2m(p, k, z3, 1.0, src, q)
    tgt = 2 * d + rng.uniform(-1, 1, (20, 3))
    ph, gr = O.h_m2p(p, k, z3, 1.0, M, tgt)
    ep, eg = _direct(k, src, q, tgt)
    assert rel(ph, ep) < tol_phi and rel(gr, eg) < tol_grad
    # P2M -> M2L -> L2P (target box at 2d; cells differ by d = target - source)
    tl = rng.uniform(-1, 1, (20, 3))
    ph, gr = O.h_l2p(p, k, z3, 1.0, O.h_m2l(p, k, d) @ M, tl)
    ep, eg = _direct(k, src - 2 * d, q, tl)
    assert rel(ph, ep) < tol_phi and rel(gr, eg) < tol_grad
    # P2L -> L2P
    ph, gr = O.h_l2p(p, k, z3, 1.0, O.h_p2l(p, k, -2.0 * d, 1.0, src, q), tl)
    ep, eg = _direct(k, src + 2 * d, q, tl)
    assert rel(ph, ep) < tol_phi and rel(gr, eg) < tol_grad
    for o in (0, 5, 7):
        off = np.array([((o >> 2) & 1) - 0.5, ((o >> 1) & 1) - 0.5, (o & 1) - 0.5])
        # child P2M (child units, k/2) -> M2M -> parent M2P (parent units, k)
        Mp = O.h_m2m(p, k, o) @ O.h_p2m(p, k / 2, z3, 1.0, src, q)
        tp = 2 * d + rng.uniform(-1, 1, (20, 3))
        ph, gr = O.h_m2p(p, k, z3, 1.0, Mp, tp)
        ep, eg = _direct(k, off + src / 2, q, tp)
        assert rel(ph, ep) < tol_phi and rel(gr, eg) < tol_grad
        # parent P2L from far charges -> L2L into the child -> L2P in child units
        far = 2 * d + rng.uniform(-1, 1, (30, 3))
        Lc = O.h_l2l(p, k, o) @ O.h_p2l(p, k, z3, 1.0, far, q)
        ph, gr = O.h_l2p(p, k / 2, z3, 1.0, Lc, tl)
        ep, eg = _direct(k / 2, (far - off) * 2, q, tl)
        assert rel(ph, ep) < tol_phi and rel(gr, eg) < tol_grad


@pytest.mark.parametrize("layout", ["uniform", "clustered"])
def test_oracle_helmholtz_fmm_matches_direct_sum(layout):
    rng = np.random.default_rng(7)
    if layout == "uniform":
        pts = rng.uniform(0, 1, (1500, 3))
    else:  # deep adaptive tree: W and X lists
        pts = np.vstack([rng.normal(0.3, 0.03, (750, 3)), rng.uniform(0, 1, (750, 3))])
    q = rng.normal(size=len(pts)) + 1j * rng.normal(size=len(pts))
    f = O.Fmm(pts, p=12, leaf_points=40)
    if layout == "clustered":
        assert f.nlevels >= 6
        assert any(len(f.list("w", l)[1]) for l in range(f.nlevels))
        assert any(len(f.list("x", l)[1]) for l in range(f.nlevels))
    for k in (25.0, 5.0 + 0.5j):
        ph, gr = f.evaluate_helmholtz(k, q)
        s = rng.choice(len(pts), 150, replace=False)
        ep, eg = _direct(k, pts, q, pts[s])
        assert rel(ph[s], ep / (4 * np.pi)) < 3e-5
        assert rel(gr[s], eg / (4 * np.pi)) < 1e-4


# ---- the product's host side against the oracle (no GPU) -----------------------------------

def test_product_tables_match_oracle_tables():
    import fmm_amd

    for p, kh in [(6, 0.8), (8, 1.3 + 0.2j), (10, 0.05), (12, 3.1), (8, 2.0j)]:
        for kind, i in [("m2m", 3), ("l2l", 6), ("m2m", 0), ("l2l", 7), ("m2l", 0), ("m2l", 157), ("m2l", 300)]:
            A = fmm_amd.helmholtz_table(p, kh, kind, i)
            if kind == "m2m":
                B = O.h_m2m(p, kh, i)
            elif kind == "l2l":
                B = O.h_l2l(p, kh, i)
            else:
                B = O.h_m2l(p, kh, O.offset_of_index(i))
            assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max(), (p, kh, kind, i)


def test_rotation_pieces_compose_the_oracle_tables():
    # the device's per-pair M2L pieces (rotation blocks per offset, coaxial blocks per
    # distance with the stored forms' scales, the transpose as the rotation back) composed
    # on the host: the oracle's M2L table of every checked offset, to rounding
    import fmm_amd

    for p, kh in [(4, 0.9), (8, 1.3 + 0.2j), (12, 3.1), (15, 2.2), (10, 0.05), (8, 2.0j)]:
        for i in (0, 37, 73, 150, 211, 315):
            A = fmm_amd.helmholtz_table(p, kh, "m2l_rcr", i)
            B = O.h_m2l(p, kh, O.offset_of_index(i))
            assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max(), (p, kh, i)


def test_degree_rule_matches_the_reference_examples():
    """fmm-exec/src/helmholtz/degree.rs:63-76: the plateau 15 at 1e-6 and kD = 100 on the
    unit cube giving 35 and 22 on levels 2 and 3."""
    import fmm_amd

    assert fmm_amd.helmholtz_degree(1.0, 0.01, 1e-6) == 15
    assert fmm_amd.helmholtz_degree(100.0, 0.125, 1e-6) == 35
    assert fmm_amd.helmholtz_degree(100.0, 0.0625, 1e-6) == 22


def test_wavenumber_rules():
    import fmm_amd

    for k in (0.0, -1.0, 1.0 - 0.5j):
        with pytest.raises(fmm_amd.FmmError):
            fmm_amd.HelmholtzFmmBuilder(k)
