"""The Helmholtz exponential (plane-wave) M2L (docs/CONVENTIONS.md §3.15 "Exponential M2L";
fmm-tables/src/helmholtz/plane_wave tests): the complex Bessel norms behind the counts
against scipy, each level's quadrature against e^{i kappa |u|}/|u| on its region (the
reference's check: within 2x FMM3D's stated error), one pair's translation against the
dense M2L of the oracle, one parent's merge and shift against its pairs, and the GPU FMM
with the exponential M2L against the oracle FMM and the direct sum."""
import numpy as np
import pytest
from scipy.special import jv

import fmm_amd
from oracle import oracle as O


@pytest.mark.parametrize("z", [3.0 + 0.5j, 20 - 2j, 0.3 + 0j, 50 + 5j, 1e-3 - 1e-3j])
def test_bessel_norms_match_scipy(z):
    a = fmm_amd.bessel_j_norms(z, 80)
    b = np.abs(jv(np.arange(81), z))
    big = b > 1e-280
    assert np.all(np.abs(a[big] - b[big]) <= 1e-12 * b[big] + 1e-300)


def _region_error(q, kappa, npts=400, seed=0):
    rng = np.random.default_rng(seed)
    z = rng.uniform(1, 4, npts)
    x = rng.uniform(-4, 4, npts)
    y = rng.uniform(-4, 4, npts)
    s = np.zeros(npts, complex)
    for k in range(len(q.lambdas)):
        n = q.alpha_counts[k]
        a = 2 * np.pi * np.arange(n) / n
        ph = np.exp(1j * q.t[k] * (x[:, None] * np.cos(a) + y[:, None] * np.sin(a))).sum(1)
        s += q.weights[k] / n * np.exp(-q.lambdas[k] * z) * ph
    r = np.sqrt(x * x + y * y + z * z)
    return np.abs(s - np.exp(1j * kappa * r) / r).max()


@pytest.mark.parametrize("kappa", [6.25, 3.125, 0.78125, 0.05, 30.0, 3.0 + 0.5j, 2.0 + 2.0j])
def test_quadrature_reproduces_the_kernel_on_its_region(kappa):
    """Both node maps (real t where Im kappa >= pi/2, the contour below) and the counts from
    the Bessel tails: the waves reproduce e^{i kappa |u|}/|u| on 1 <= z <= 4, |x|, |y| <= 4
    within twice FMM3D's stated error (a mistyped node, weight or map fails by orders)."""
    q = fmm_amd.helmholtz_plane_wave_quadrature(kappa)
    assert q.domain > 0 and q.stated_error < 1e-6
    assert np.all(q.alpha_counts >= 1) and np.all(q.fourier_counts >= 0)
    assert np.allclose(q.t ** 2, q.lambdas ** 2 + kappa ** 2)
    assert _region_error(q, kappa) <= 2 * q.stated_error


def test_domains():
    assert fmm_amd.helmholtz_plane_wave_quadrature(0.01).domain == 1
    assert fmm_amd.helmholtz_plane_wave_quadrature(0.5).domain == 6
    assert fmm_amd.helmholtz_plane_wave_quadrature(np.pi).domain == 9  # on a breakpoint: the one above
    assert fmm_amd.helmholtz_plane_wave_quadrature(0.01 + 0.03j).domain == 1 + 23  # Im in [0.02, 0.04)
    assert fmm_amd.helmholtz_plane_wave_quadrature(16 * np.pi).domain == 0
    assert fmm_amd.helmholtz_plane_wave_quadrature(1.0 + 12 * np.pi * 1j).domain == 0


@pytest.mark.parametrize("level", [2, 3, 5])
def test_pair_translation_within_the_stated_error(level):
    """Multipoles of complex charges in the source box translated by the waves and by the
    oracle's dense M2L: their potentials in the target box differ by less than the stated
    error per unit sum |q| on every sampled V offset (p = 15, k = 25 levels of the unit cube)."""
    p, kh = 15, 25.0 / 2 ** level / 2
    q0 = fmm_amd.helmholtz_plane_wave_quadrature(2 * kh)
    rng = np.random.default_rng(level)
    worst = 0.0
    for t in range(0, 316, 11):
        d = np.array(O.offset_of_index(t))
        src = rng.uniform(-1, 1, (20, 3))
        q = rng.normal(size=20) + 1j * rng.normal(size=20)
        M = O.h_p2m(p, kh, np.zeros(3), 1.0, src, q)
        Ld = O.h_m2l(p, kh, d) @ M
        Lp = fmm_amd.helmholtz_plane_wave_m2l(p, kh, t, M)
        tl = rng.uniform(-1, 1, (10, 3))
        pd, _ = O.h_l2p(p, kh, np.zeros(3), 1.0, Ld, tl)
        pp, _ = O.h_l2p(p, kh, np.zeros(3), 1.0, Lp, tl)
        worst = max(worst, np.abs(pp - pd).max() / np.abs(q).sum())
    assert worst < q0.stated_error, worst


def test_merge_equals_its_pairs():
    """One parent's merge and shift equals its pairs translated one by one along the merge's
    directions to rounding."""
    p, kh = 8, 25.0 / 16
    rng = np.random.default_rng(4)
    n = (p + 1) ** 2
    present = rng.uniform(size=216) < 0.6
    for a in range(216):
        if all(2 <= x <= 3 for x in (a // 36, (a // 6) % 6, a % 6)):
            present[a] = False
    mults = np.zeros((216, n), complex)
    for a in np.flatnonzero(present):
        mults[a] = O.h_p2m(p, kh, np.zeros(3), 1.0, rng.uniform(-1, 1, (4, 3)),
                           rng.normal(size=4) + 1j * rng.normal(size=4))
    merged = fmm_amd.helmholtz_plane_wave_merge(p, kh, mults, present)
    pairs = np.zeros((8, n), complex)
    for a in np.flatnonzero(present):
        av = (a // 36, (a // 6) % 6, a % 6)
        for c in range(8):
            r = fmm_amd.plane_wave_slot_direction(int(a), c)
            if r < 0:
                continue
            d = tuple(((c >> (2 - i)) & 1) + 2 - av[i] for i in range(3))
            pairs[c] += fmm_amd.helmholtz_plane_wave_m2l(p, kh, O.offset_index(d), mults[a], direction=r)
    assert np.abs(merged - pairs).max() <= 1e-11 * np.abs(pairs).max()
