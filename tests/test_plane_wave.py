"""The plane-wave M2L (docs/CONVENTIONS.md:1305-1434): an approximate operator beside the
exact M2L, checked as the reference checks its own (fmm-tables/src/plane_wave/tests.rs):
FMM3D's quadrature tables against their known sums and stated errors, one pair's
translation against the dense M2L within the quadrature's measured error, the merge-and-
shift geometry against its invariants, one parent's merge and shift against the sum of its
pairs, and the GPU FMM with plane-wave M2L against the oracle FMM (exact M2L) and the
direct sum."""
import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O


def measured_error(q):
    """The quadrature's largest error against 1/|u| (box width 1) with its even alpha
    counts, on the grid z in {1, 1.25, .., 4}, x, y in {-4, -3.5, .., 4}
    (Quadrature::measured_error, fmm-tables/src/plane_wave/quadrature.rs:108-137)."""
    even = q.alpha_nodes + q.alpha_nodes % 2
    z = 1.0 + 0.25 * np.arange(13)
    x = -4.0 + 0.5 * np.arange(17)
    Z, X, Y = np.meshgrid(z, x, x, indexing="ij")
    s = np.zeros_like(Z)
    for k, n in enumerate(even):
        a = 2 * np.pi * np.arange(n) / n
        inner = np.cos(q.lambdas[k] * (X[..., None] * np.cos(a) + Y[..., None] * np.sin(a))).sum(-1)
        s += q.weights[k] / n * np.exp(-q.lambdas[k] * Z) * inner
    return np.abs(s - 1 / np.sqrt(X * X + Y * Y + Z * Z)).max()


def test_quadrature_tables_are_fmm3ds():
    for s in range(2, 40):
        q = fmm_amd.plane_wave_quadrature(0, s)
        assert len(q.lambdas) == len(q.weights) == len(q.alpha_nodes) == s
        assert np.all(np.diff(q.lambdas) > 0) and 0 < q.max_error < 0.2
        assert q.half_terms == int(np.sum(q.alpha_nodes + q.alpha_nodes % 2)) // 2
    # FMM3D's sums of the alpha counts for the sizes its lfmm3d chooses (the reference's
    # quadrature_tables_are_consistent)
    assert [int(fmm_amd.plane_wave_quadrature(0, s).alpha_nodes.sum()) for s in (12, 20, 29, 37)] == [217, 705, 1627,
                                                                                                   2770]
    counts = [len(fmm_amd.plane_wave_quadrature(p).lambdas) for p in range(21)]
    assert counts[12] == 21 and all(b >= a for a, b in zip(counts, counts[1:]))
    q = fmm_amd.plane_wave_quadrature(12)
    assert q.half_terms == 399 and q.fourier.max() == 13
    assert np.all(fmm_amd.plane_wave_quadrature(3).fourier <= 4)
    with pytest.raises(fmm_amd.FmmError):
        fmm_amd.plane_wave_quadrature(21)
    with pytest.raises(fmm_amd.FmmError):
        fmm_amd.plane_wave_quadrature(6, 40)


@pytest.mark.parametrize("s", [8, 12])
def test_stated_and_measured_errors(s):
    """With many alpha nodes the lambda sum meets FMM3D's stated error; with FMM3D's alpha
    counts the error over the region is 5-25x that (the reference's stated_and_measured_errors)."""
    q = fmm_amd.plane_wave_quadrature(0, s)
    worst = 0.0
    a = 2 * np.pi * np.arange(400) / 400
    for z in (1.0, 1.3, 2.0, 3.1, 4.0):
        for x, y in ((0.0, 0.0), (1.0, -2.0), (4.0, 4.0), (-3.5, 2.5)):
            inner = np.cos(q.lambdas[:, None] * (x * np.cos(a) + y * np.sin(a))[None, :]).sum(1)
            approx = (q.weights / 400 * np.exp(-q.lambdas * z) * inner).sum()
            worst = max(worst, abs(approx - 1 / np.sqrt(x * x + y * y + z * z)))
    assert worst <= 1.05 * q.max_error, worst
    m = measured_error(q)
    assert 5 * q.max_error < m < 25 * q.max_error, m / q.max_error


def _pair_errors(p, nodes, step=5, seed=3):
    rng = np.random.default_rng(seed)
    _, _, m2l = fmm_amd.tables(p)
    worst = 0.0
    for t in range(0, 316, step):
        d = np.array(O.offset_of_index(t))
        src = rng.uniform(-1, 1, (30, 3))
        q = rng.normal(size=30)
        M = O.p2m(p, np.zeros(3), 1.0, src, q)
        Ld = m2l[t] @ M
        Lp = fmm_amd.plane_wave_m2l(p, t, M, nodes=nodes)
        tg = rng.uniform(-1, 1, (20, 3))
        pd, _ = O.l2p(p, np.zeros(3), 1.0, Ld, tg)
        pp, _ = O.l2p(p, np.zeros(3), 1.0, Lp, tg)
        worst = max(worst, np.abs(pp - pd).max() / np.abs(q).sum())
    return worst


@pytest.mark.parametrize("p", [6, 8, 12])
def test_pair_translation_within_the_quadrature_error(p):
    """Multipoles of charges in the source box translated by the plane waves and by the
    dense table: their potentials in the target box differ by at most the quadrature's
    measured error per unit sum |q| (the reference's pair_m2l_within_the_quadrature_error;
    scaled coordinates have box width 2, so 1/|x| is half the box-width potential)."""
    q = fmm_amd.plane_wave_quadrature(p)
    assert _pair_errors(p, 0) <= 0.5 * measured_error(q)


def test_pair_translation_is_exact_with_the_finest_quadrature():
    """With FMM3D's 39-node table the quadrature error falls below 1e-12 and the plane-wave
    steps give the dense M2L to that level: a dropped term, sign or frame would not."""
    assert _pair_errors(12, 39, step=3) < 2e-12


def test_merge_geometry():
    """Every pair slot -> child that is not adjacent has a direction along whose axis the
    child lies 2 or 3 boxes from the slot, laterally within the quadrature's region; the
    directions put the 1512 pairs into 248 (slot, direction) merges whose shifts to the
    parent's centre take 104 distinct values (the reference's tables_have_their_sizes)."""
    frames = {0: lambda d: d, 1: lambda d: (d[0], -d[2], d[1]), 2: lambda d: (-d[2], d[1], d[0])}
    pairs, merges, shifts = 0, set(), set()
    for slot in range(216):
        a = (slot // 36, (slot // 6) % 6, slot % 6)
        for c in range(8):
            b = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
            d = tuple(b[i] + 2 - a[i] for i in range(3))
            r = fmm_amd.plane_wave_slot_direction(slot, c)
            if max(abs(x) for x in d) <= 1:
                assert r == -1
                continue
            assert 0 <= r < 6
            pairs += 1
            dp = frames[r // 2](d)
            along = dp[2] if r % 2 == 0 else -dp[2]
            assert along in (2, 3) and dp[0] ** 2 + dp[1] ** 2 <= 18
            merges.add((slot, r))
            sp = frames[r // 2](tuple(2.5 - x for x in a))
            shifts.add((r % 2,) + tuple(sp))
    assert pairs == 1512 and len(merges) == 248 and len(shifts) == 104
    assert fmm_amd.plane_wave_slot_direction(216, 0) == -2


@pytest.mark.parametrize("p", [6, 10])
def test_merge_equals_its_pairs(p):
    """One parent's merge and shift equals the sum of its pairs translated one by one in the
    same directions to rounding, and the dense M2L within the quadrature's error."""
    rng = np.random.default_rng(p)
    nc = (p + 1) ** 2
    present = rng.uniform(size=216) < 0.7
    for a in range(216):
        if all(2 <= x <= 3 for x in (a // 36, (a // 6) % 6, a % 6)):
            present[a] = False  # the parent's own children: adjacent to every child
    mults = np.zeros((216, nc))
    for a in np.flatnonzero(present):
        mults[a] = O.p2m(p, np.zeros(3), 1.0, rng.uniform(-1, 1, (6, 3)), rng.normal(size=6))
    merged = fmm_amd.plane_wave_merge(p, mults, present)
    _, _, m2l = fmm_amd.tables(p)
    pairs = np.zeros((8, nc))
    dense = np.zeros((8, nc))
    for a in np.flatnonzero(present):
        av = (a // 36, (a // 6) % 6, a % 6)
        for c in range(8):
            r = fmm_amd.plane_wave_slot_direction(int(a), c)
            if r < 0:
                continue
            d = tuple(((c >> (2 - i)) & 1) + 2 - av[i] for i in range(3))
            t = O.offset_index(d)
            pairs[c] += fmm_amd.plane_wave_m2l(p, t, mults[a], direction=r)
            dense[c] += m2l[t] @ mults[a]
    assert np.abs(merged - pairs).max() <= 1e-12 * np.abs(pairs).max()
    tg = np.random.default_rng(9).uniform(-1, 1, (10, 3))
    scale = max(np.abs(O.l2p(p, np.zeros(3), 1.0, dense[c], tg)[0]).max() for c in range(8))
    for c in range(8):
        pm, _ = O.l2p(p, np.zeros(3), 1.0, merged[c], tg)
        pd, _ = O.l2p(p, np.zeros(3), 1.0, dense[c], tg)
        assert np.abs(pm - pd).max() <= 100 * measured_error(fmm_amd.plane_wave_quadrature(p)) * scale


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,p,leaf", [("cube", 8000, 8, 40), ("plummer", 8000, 10, 32), ("cube", 20000, 12, 40)])
def test_plane_wave_fmm_matches_oracle(kind, n, p, leaf):
    """The GPU FMM with the merge-and-shift plane-wave M2L against the oracle FMM (exact
    M2L) and the direct sum: its error stays within 1.5x that of the exact M2L (the
    reference picks the quadrature so that it errs at most a tenth of the truncation), and
    it differs from the oracle by less than the truncation error."""
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
    sample = np.arange(0, n, 37)
    dphi, _ = O.direct_sum(pts, q, pts[sample])
    dphi /= 4 * np.pi  # the FMM's kernel is 1 / (4 pi r)
    e_dense = np.linalg.norm(b.potential - ophi) / np.linalg.norm(ophi)
    t_dense = np.linalg.norm(b.potential[sample] - dphi) / np.linalg.norm(dphi)
    t_pw = np.linalg.norm(a.potential[sample] - dphi) / np.linalg.norm(dphi)
    e_pw = np.linalg.norm(a.potential - ophi) / np.linalg.norm(ophi)
    eg = np.linalg.norm(a.gradient - ograd) / np.linalg.norm(ograd)
    print(f"  plane wave vs oracle {e_pw:.2e} (grad {eg:.2e}); vs direct {t_pw:.2e}, dense vs direct {t_dense:.2e}",
          flush=True)
    assert e_dense < 1e-12
    assert t_pw <= 1.5 * t_dense
    assert e_pw <= t_dense


@pytest.mark.gpu
def test_plane_wave_in_the_mixed_operator():
    """The mixed operator with the plane-wave M2L (f64 waves beside the f32 P2P, as the
    reference runs its plane-wave tables in either precision): within the mixed bar of the
    f64 plane-wave operator on the same plan."""
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    print("case plane-wave fmm mixed", flush=True)
    pts = O.Rng(11).cube(12000)
    q = O.Rng(5).normals(12000)
    f64 = fmm_amd.FmmBuilder(8).max_points_per_leaf(40).m2l("plane_wave").build(pts)
    mixed = fmm_amd.Fmm(f64.plan, fmm_amd.Operator(f64.plan, 8, "mixed", m2l="plane_wave"))
    a, b = f64.evaluate(q), mixed.evaluate(q)
    e = np.linalg.norm(b.potential - a.potential) / np.linalg.norm(a.potential)
    eg = np.linalg.norm(b.gradient - a.gradient) / np.linalg.norm(a.gradient)
    print(f"  mixed vs f64 plane wave: {e:.2e} / {eg:.2e}", flush=True)
    assert e < 1e-6 and eg < 1e-6
