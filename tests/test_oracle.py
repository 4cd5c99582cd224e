"""Pins for the oracle (the f64 CPU restatement the GPU path is checked against).

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a
wrong sign, a transposed index) fails: published sequences and known answers of the
reference's own tests, its committed mpmath fixtures, closed forms and identities, a
brute-force geometric list oracle, and the reference's published whole-FMM errors.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O


def idx(n, m):
    return n * n + n + m


# ---- rng, P2P, direct sum ---------------------------------------------------------


def test_splitmix64_published_sequence():
    # fmm-validate/src/rng.rs:54-60 (Vigna's splitmix64.c outputs for seed 0)
    r = O.Rng(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_points_are_deterministic_and_inside():
    a = O.Rng(11).cube(200, (0.3, -1.2, 2.5), 0.7)
    b = O.Rng(11).cube(200, (0.3, -1.2, 2.5), 0.7)
    assert np.array_equal(a, b)
    assert np.all(np.abs(a - [0.3, -1.2, 2.5]) <= 0.7 + 1e-15)
    s = O.Rng(3).sphere(500, (0, 0, 0), 2.0)
    assert np.allclose(np.linalg.norm(s, axis=1), 2.0, rtol=1e-15, atol=0)
    pl = O.Rng(4).plummer(2000, (0, 0, 0), 0.1)
    assert np.max(np.linalg.norm(pl, axis=1)) <= 1.0 + 1e-12
    g = O.Rng(5).gaussian_clusters(1001, [[0, 0, 0], [1, 1, 1]], 0.01)
    assert len(g) == 1001 and np.max(np.linalg.norm(g[:501], axis=1)) <= 0.04 + 1e-12
    q = O.Rng(7).normals(20000)
    assert abs(q.mean()) < 0.03 and abs(q.std() - 1) < 0.03


def test_p2p_known_answers():
    # fmm-ref/src/p2p.rs:51-60
    pts = np.array([[0.0, 0, 0], [3, 4, 0]])
    q = np.array([2.0, -1.0])
    phi, grad = O.p2p(pts, q, pts)
    assert list(phi) == [-1.0 / 5.0, 2.0 / 5.0]
    assert list(grad[1]) == [-2.0 * 3.0 / 125.0, -2.0 * 4.0 / 125.0, 0.0]
    ophi, _ = O.direct_sum(pts, q, pts)
    assert list(ophi) == list(phi)
    # fmm-ref/src/p2p.rs:495-503: term (3, 0, 4), q = 2
    phi, grad = O.p2p(np.zeros((1, 3)), np.array([2.0]), np.array([[3.0, 0, 4]]))
    assert phi[0] == 0.4 and list(grad[0]) == [-6.0 / 125.0, 0.0, -8.0 / 125.0]


def test_neumaier_and_compensated_direct_sum():
    # fmm-ref/src/p2p.rs:463-486
    assert O.lib().orc_neumaier_sum(np.array([1.0, 1e100, 1.0, -1e100]), 4) == 2.0
    src = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [-1.0, 0, 0]])
    q = np.array([1.0, 1e100, 1.0, -1e100])
    comp, _ = O.direct_sum(src, q, np.zeros((1, 3)), gradients=False)
    plain, _ = O.p2p(src, q, np.zeros((1, 3)), gradients=False)
    assert comp[0] == 2.0 and plain[0] == 0.0


def test_error_norms():
    # fmm-validate/src/metrics.rs:26-34
    l2, mx = O.error_norms([3.0, 4.5], [3.0, 4.0])
    assert l2 == pytest.approx(0.5 / 5.0, rel=1e-15) and mx == pytest.approx(0.5 / 4.0, rel=1e-15)


# ---- harmonics against the reference's mpmath fixtures -----------------------------


def _check(ours, ref, p, rtol):
    # per degree, relative to the largest magnitude of that degree
    for n in range(p + 1):
        sl = slice(n * n, (n + 1) * (n + 1))
        scale = max(np.max(np.abs(ref[sl])), 1e-300)
        assert np.max(np.abs(ours[sl] - ref[sl])) <= rtol * scale, (n, ours[sl], ref[sl])


@pytest.mark.parametrize("tag", ["A", "B", "C"])
def test_harmonics_match_mpmath_fixtures(golden_dir, tag):
    g = np.load(os.path.join(golden_dir, "harmonics.npz"))
    p = int(g[f"{tag}_p"])
    for kind, fn, gfn in (("regular", O.regular, O.regular_grad), ("irregular", O.irregular, O.irregular_grad)):
        if f"{tag}_{kind}_x" not in g:
            continue
        xs, vals = g[f"{tag}_{kind}_x"], g[f"{tag}_{kind}_value"]
        grads = g[f"{tag}_{kind}_grad"] if f"{tag}_{kind}_grad" in g else None
        for i, x in enumerate(xs):
            _check(fn(p, x), vals[i], p, 5e-14 if p <= 20 else 2e-13)
            if grads is not None:
                ours = gfn(p, x)
                for k in range(3):
                    _check(ours[k], grads[i][k], p, 5e-14)


def test_harmonic_examples_and_laplace():
    # CONVENTIONS §3.3 examples: R_1^0 = z, R_1^1 = (x + iy)/2, R_2^0 = (2z^2 - x^2 - y^2)/4
    x = np.array([0.3, -0.7, 0.4])
    r = O.regular(2, x)
    assert r[idx(1, 0)] == pytest.approx(0.4)
    assert r[idx(1, 1)] == pytest.approx(0.15) and r[idx(1, -1)] == pytest.approx(-0.35)
    assert r[idx(2, 0)] == pytest.approx((2 * 0.16 - 0.09 - 0.49) / 4)
    # harmonicity: the Laplacian by central differences vanishes
    h = 1e-3
    for fn, pt in ((O.regular, x), (O.irregular, np.array([2.1, -1.3, 2.7]))):
        lap = -6 * fn(6, pt)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            lap += fn(6, pt + e) + fn(6, pt - e)
        assert np.max(np.abs(lap / h**2)) < 1e-4 * np.max(np.abs(fn(6, pt)))


# ---- leaf operators and translations ---------------------------------------------------


def _cloud(seed, n, c, r):
    rng = O.Rng(seed)
    return rng.ball(n, c, r), rng.charges(n)


def test_p2m_m2p_and_p2l_l2p_converge_to_direct():
    p = 14
    src, q = _cloud(1, 40, (0, 0, 0), 0.9)
    far = O.Rng(2).sphere(20, (0, 0, 0), 5.0)
    exact, egrad = O.direct_sum(src, q, far)
    m = O.p2m(p, (0, 0, 0), 1.0, src, q)
    phi, grad = O.m2p(p, (0, 0, 0), 1.0, m, far)
    assert np.max(np.abs(phi - exact)) < 1e-9 * np.max(np.abs(exact))
    assert np.max(np.abs(grad - egrad)) < 1e-8 * np.max(np.abs(egrad))
    # local expansion of far sources, evaluated near the centre (frame radius 0.5)
    fsrc, fq = O.Rng(3).sphere(30, (0, 0, 0), 6.0), O.Rng(4).charges(30)
    tgt = O.Rng(5).ball(20, (0, 0, 0), 0.5)
    loc = O.p2l(p, (0, 0, 0), 0.5, fsrc, fq)
    phi, grad = O.l2p(p, (0, 0, 0), 0.5, loc, tgt)
    ex, eg = O.direct_sum(fsrc, fq, tgt)
    assert np.max(np.abs(phi - ex)) < 1e-10 * np.max(np.abs(ex))
    assert np.max(np.abs(grad - eg)) < 1e-9 * np.max(np.abs(eg))


def test_m2m_is_exact():
    # CONVENTIONS §3.11: M2M of a degree-p P2M equals P2M in the output frame
    p = 10
    c, r = np.array([0.25, -0.25, 0.25]), 0.25
    src, q = _cloud(6, 25, c, 0.4)
    m_child = O.p2m(p, c, r, src, q)
    b = (c - 0) / 0.5
    m_parent = O.m2m(p, b, r / 0.5, m_child)
    direct = O.p2m(p, (0, 0, 0), 0.5, src, q)
    assert np.max(np.abs(m_parent - direct)) < 1e-13 * np.max(np.abs(direct))


def test_l2l_preserves_the_potential():
    p = 10
    fsrc, fq = O.Rng(7).sphere(30, (0, 0, 0), 8.0), O.Rng(8).charges(30)
    loc = O.p2l(p, (0, 0, 0), 1.0, fsrc, fq)
    cc, rc = np.array([-0.5, 0.5, 0.5]), 0.5
    child = O.l2l(p, (cc - 0) / 1.0, rc / 1.0, loc)
    tgt = O.Rng(9).ball(15, cc, 0.4)
    a, ga = O.l2p(p, (0, 0, 0), 1.0, loc, tgt)
    b, gb = O.l2p(p, cc, rc, child, tgt)
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))
    assert np.max(np.abs(ga - gb)) < 1e-12 * np.max(np.abs(ga))


def test_m2l_converges_to_p2l():
    # M2L of P2M equals P2L in the output frame up to the truncation of §3.11
    src, q = _cloud(10, 30, (0, 0, 0), 0.8)
    d = np.array([2.0, -2.0, 3.0])
    errs = []
    for p in (4, 8, 12):
        m = O.p2m(p, (0, 0, 0), 1.0, src, q)
        l = O.m2l(p, 2 * d, 1.0, m)
        ref = O.p2l(p, 2 * d, 1.0, src, q)
        errs.append(np.max(np.abs(l - ref)) / np.max(np.abs(ref)))
    assert errs[2] < errs[1] < errs[0] and errs[2] < 1e-7


def test_offsets_and_tables():
    seen = []
    for x in range(-3, 4):
        for y in range(-3, 4):
            for z in range(-3, 4):
                if max(abs(x), abs(y), abs(z)) >= 2:
                    seen.append((x, y, z))
    assert len(seen) == 316
    for i, d in enumerate(seen):
        assert O.offset_index(d) == i and O.offset_of_index(i) == d
    assert O.offset_index((-1, 0, 1)) == -1
    p = 5
    m2m, l2l, m2l = O.tables(p)
    n = (p + 1) ** 2
    x = O.Rng(11).charges(n)
    # table columns equal the direct operators (octant 4 = (+1, -1, -1))
    b = 0.5 * np.array([1.0, -1.0, -1.0])
    assert np.allclose(m2m[4] @ x, O.m2m(p, b, 0.5, x), rtol=0, atol=1e-14)
    assert np.allclose(l2l[4] @ x, O.l2l(p, b, 0.5, x), rtol=0, atol=1e-14)
    i = O.offset_index((2, 0, -3))
    assert np.allclose(m2l[i] @ x, O.m2l(p, 2 * np.array([2.0, 0, -3]), 1.0, x), rtol=0, atol=1e-14)
    # inversion: M2L(-d) = diag((-1)^j) M2L(d) diag((-1)^n)   (CONVENTIONS §3.12)
    deg = np.array([int(np.floor(np.sqrt(k))) for k in range(n)])
    sgn = (-1.0) ** deg
    for d in ((2, 0, -3), (3, 3, 3), (0, 2, 1)):
        a = m2l[O.offset_index(d)]
        b = m2l[O.offset_index(tuple(-v for v in d))]
        assert np.allclose(b, sgn[:, None] * a * sgn[None, :], rtol=1e-13, atol=1e-300)


# ---- tree and lists --------------------------------------------------------------------


def _level2_grid():
    c = (np.arange(4) + 0.5) / 4
    return np.array([[x, y, z] for x in c for y in c for z in c])


def _box_lookup(f):
    keys = {}
    for l in range(f.nlevels):
        for b, k in enumerate(f.box_keys(l)):
            keys[int(k)] = (l, b)
    return keys


def test_uniform_level_two_tree_matches_hand_counts():
    # fmm-plan/src/interaction_manager_tests.rs:251-285
    f = O.Fmm(_level2_grid(), p=2, leaf_points=1)
    assert f.nlevels == 3 and len(f.leaf_keys()) == 64
    keys = [int(k) for k in f.box_keys(2)]
    near_off, _ = f.list("near", 2)
    v_off, _, _ = f.list("v", 2)
    lk = [int(k) for k in f.leaf_keys()]
    for index, u_count, v_count in (((0, 0, 0), 7, 56), ((1, 0, 0), 11, 52), ((1, 1, 0), 17, 46), ((1, 1, 1), 26, 37)):
        k = O.key(2, *index)
        r = lk.index(k)
        b = keys.index(k)
        assert near_off[r + 1] - near_off[r] == u_count + 1
        assert v_off[b + 1] - v_off[b] == v_count
    w_off, _ = f.list("w", 2)
    x_off, _ = f.list("x", 2)
    assert w_off[-1] == 0 and x_off[-1] == 0


def test_one_refined_octant_gives_expected_w_and_x():
    # fmm-plan/src/interaction_manager_tests.rs:287-356
    pts = [[0.0, 0.0, 0.0]]
    for i in range(2):
        for j in range(2):
            for k in range(2):
                if (i, j, k) != (0, 0, 0):
                    pts.append([0.125 + 0.25 * i, 0.125 + 0.25 * j, 0.125 + 0.25 * k])
                if (i, j, k) != (0, 0, 0):
                    pts.append([float(i), float(j), float(k)])
    f = O.Fmm(np.array(pts), p=2, leaf_points=1)
    lk = [int(k) for k in f.leaf_keys()]
    assert len(lk) == 15
    child = lambda ix: O.key(2, *ix)
    octant = lambda ix: O.key(1, *ix)
    keys2 = [int(k) for k in f.box_keys(2)]
    w_off, w_ent = f.list("w", 1)
    lvl1 = [k for k in lk if k & 0x7FFF == 1]

    def w_of(k):
        r = lvl1.index(k)
        return sorted(keys2[e] for e in w_ent[w_off[r]:w_off[r + 1]])

    assert w_of(octant((1, 0, 0))) == sorted(child((0, y, z)) for y in range(2) for z in range(2))
    assert w_of(octant((1, 1, 1))) == sorted(child((a, b, c)) for a in range(2) for b in range(2) for c in range(2)
                                             if (a, b, c) != (1, 1, 1))
    x_off, x_ent = f.list("x", 2)

    def x_of(k):
        b = keys2.index(k)
        return sorted(lk[e] for e in x_ent[x_off[b]:x_off[b + 1]])

    others = sorted(octant((a, b, c)) for a in range(2) for b in range(2) for c in range(2) if (a, b, c) != (0, 0, 0))
    assert x_of(child((0, 0, 0))) == others
    assert x_of(child((1, 1, 1))) == []
    assert x_of(child((1, 0, 0))) == [k for k in others if k != octant((1, 0, 0))]
    near_off, near_ent = f.list("near", 2)
    lvl2 = [k for k in lk if k & 0x7FFF == 2]
    r = lvl2.index(child((1, 0, 0)))
    got = sorted(lk[e] for e in near_ent[near_off[r]:near_off[r + 1]])
    want = sorted([child((a, b, c)) for a in range(2) for b in range(2) for c in range(2)] + [octant((1, 0, 0))])
    assert got == want


def brute_force_lists(f):
    """The list definitions of fmm-plan/src/interaction_manager.rs:10-28, applied by
    brute force over every pair of boxes with the geometric adjacency test."""
    allk = []
    for l in range(f.nlevels):
        allk += [int(k) for k in f.box_keys(l)]
    leaves = set(int(k) for k in f.leaf_keys())
    lvl = lambda k: k & 0x7FFF
    par = lambda k: int(O.lib().orc_parent(k))
    adj = O.adjacent
    out = {}
    for b in allk:
        u = sorted(k for k in leaves if adj(k, b)) if b in leaves else []
        v, x = [], []
        if lvl(b) > 0:
            P = par(b)
            v = [k for k in allk if lvl(k) == lvl(b) and adj(par(k), P) and not adj(k, b)]
            x = [k for k in leaves if lvl(k) == lvl(P) and adj(k, P) and not adj(k, b)]
        w = []
        if b in leaves:
            w = [k for k in allk if lvl(k) > lvl(b) and adj(par(k), b) and not adj(k, b)]
        out[b] = (sorted(u), sorted(v), sorted(w), sorted(x))
    return out


@pytest.mark.parametrize("dist", ["plummer", "sphere"])
def test_lists_match_brute_force_on_adaptive_trees(dist):
    rng = O.Rng(17)
    pts = rng.plummer(600, (0.5, 0.5, 0.5), 0.05) if dist == "plummer" else rng.sphere(600, (0, 0, 0), 1.0)
    f = O.Fmm(pts, p=2, leaf_points=6)
    ref = brute_force_lists(f)
    lk = [int(k) for k in f.leaf_keys()]
    populated = [0, 0, 0, 0]
    for l in range(f.nlevels):
        keys = [int(k) for k in f.box_keys(l)]
        lvl_leaves = [k for k in lk if k & 0x7FFF == l]
        off, ent = f.list("near", l)
        woff, went = f.list("w", l)
        for r, k in enumerate(lvl_leaves):
            near = sorted(lk[e] for e in ent[off[r]:off[r + 1]])
            assert near == sorted(ref[k][0] + [k])
            # near rows ascend in (level, key), that is in leaf index
            assert list(ent[off[r]:off[r + 1]]) == sorted(ent[off[r]:off[r + 1]])
            w = sorted(int(f.box_keys(l + 1)[e]) for e in went[woff[r]:woff[r + 1]]) if woff[r + 1] > woff[r] else []
            assert w == ref[k][2]
            populated[0] += len(near) > 1
            populated[2] += len(w) > 0
        voff, vent, vaux = f.list("v", l)
        xoff, xent = f.list("x", l)
        for b, k in enumerate(keys):
            v = sorted(keys[e] for e in vent[voff[b]:voff[b + 1]])
            assert v == ref[k][1]
            assert list(vaux[voff[b]:voff[b + 1]]) == sorted(vaux[voff[b]:voff[b + 1]])
            for e, d in zip(vent[voff[b]:voff[b + 1]], vaux[voff[b]:voff[b + 1]]):
                it, js = O.decode(k), O.decode(keys[e])
                assert O.offset_index(tuple(it[a] - js[a] for a in (1, 2, 3))) == d
            x = sorted(lk[e] for e in xent[xoff[b]:xoff[b + 1]])
            assert x == ref[k][3]
            populated[1] += len(v) > 0
            populated[3] += len(x) > 0
    assert all(c > 0 for c in populated), populated


def test_tree_is_complete_and_balanced():
    rng = O.Rng(23)
    pts = np.vstack([rng.plummer(3000, (0.5, 0.5, 0.5), 0.02), rng.cube(500)])
    f = O.Fmm(pts, p=2, leaf_points=10)
    lk = [int(k) for k in f.leaf_keys()]
    # complete: leaf volumes add up to the domain
    assert sum(8.0 ** -(k & 0x7FFF) for k in lk) == pytest.approx(1.0, rel=0, abs=1e-12)
    # 2:1: adjacent leaves differ by at most one level
    by_level = {}
    for k in lk:
        by_level.setdefault(k & 0x7FFF, []).append(k)
    for a in lk:
        for b in lk:
            if abs((a & 0x7FFF) - (b & 0x7FFF)) > 1:
                assert not O.adjacent(a, b)
    # every point lies in its leaf
    leaves = f.point_leaves()
    box, side = f.domain()
    for i in range(0, len(pts), 97):
        lv, ix, iy, iz = O.decode(lk[leaves[i]])
        h = side / 2**lv
        for a, ia in zip(range(3), (ix, iy, iz)):
            assert box[a] + ia * h - 1e-12 <= pts[i, a] <= box[a] + (ia + 1) * h + 1e-12


# ---- whole FMM against the reference's published errors ----------------------------------


@pytest.mark.parametrize("p", [3, 6])
def test_fmm_reproduces_published_errors(golden_dir, p):
    pub = json.load(open(os.path.join(golden_dir, "published.json")))
    row = [r for r in pub["errors"] if r["n"] == 100000 and r["p"] == p][0]
    pts, q, sample = O.bench_problem(100000, 1000)
    f = O.Fmm(pts, p=p, leaf_points=64)
    assert len(f.leaf_keys()) == 4096
    phi, grad = f.evaluate(q)
    ephi, egrad = O.direct_sum(pts, q, pts[sample])
    ephi /= 4 * np.pi
    egrad /= 4 * np.pi
    e_phi = O.error_norms(phi[sample], ephi)[0]
    e_grad = O.error_norms(grad[sample], egrad, True)[0]
    # published to three significant digits
    assert abs(e_phi - row["phi"]) <= 0.005 * row["phi"], e_phi
    assert abs(e_grad - row["grad"]) <= 0.005 * row["grad"], e_grad


def test_fmm_adaptive_matches_direct_sum():
    rng = O.Rng(31)
    pts = np.vstack([rng.plummer(4000, (0.5, 0.5, 0.5), 0.05), rng.cube(2000)])
    q = rng.charges(len(pts))
    f = O.Fmm(pts, p=8, leaf_points=20)
    phi, grad = f.evaluate(q)
    ephi, egrad = O.direct_sum(pts, q, pts)
    ephi /= 4 * np.pi
    egrad /= 4 * np.pi
    assert O.error_norms(phi, ephi)[0] < 1e-4
    assert O.error_norms(grad, egrad, True)[0] < 1e-3
    # sources != targets
    tgt = rng.cube(300)
    f2 = O.Fmm(pts, tgt, p=8, leaf_points=20)
    phi2, grad2 = f2.evaluate(q)
    e2, g2 = O.direct_sum(pts, q, tgt)
    assert O.error_norms(phi2, e2 / (4 * np.pi))[0] < 1e-4


def test_helmholtz_direct_sum_matches_fmm3d(golden_dir):
    # the reference's fixture of FMM3D's direct sum (fmm-ref/fixtures/helmholtz_fmm3d.json):
    # real, complex and imaginary k, real and complex charges
    z = np.load(os.path.join(golden_dir, "helmholtz_fmm3d.npz"))
    src, tgt = z["sources"], z["targets"]
    pts = np.vstack([src, tgt])
    for c in range(len(z["k"])):
        q = z["charges_real"] if z["charge_kind"][c] == 0 else z["charges_complex"]
        pot, grad = O.helmholtz_direct_sum(z["k"][c], src, q, pts)
        pot /= 4 * np.pi
        grad /= 4 * np.pi
        ep = np.linalg.norm(pot - z["potential"][c]) / np.linalg.norm(z["potential"][c])
        eg = np.linalg.norm(grad - z["gradient"][c]) / np.linalg.norm(z["gradient"][c])
        assert ep < 1e-13 and eg < 1e-13, (c, z["k"][c], ep, eg)
    # k = 0 is Laplace's kernel (potential bit for bit, gradient to rounding)
    pot0, grad0 = O.helmholtz_direct_sum(0.0, src, z["charges_real"], pts)
    lp, lg = O.direct_sum(src, z["charges_real"].real, pts)
    assert np.array_equal(pot0.real, lp) and not np.any(pot0.imag)
    assert np.abs(grad0.real - lg).max() <= 1e-14 * np.abs(lg).max() and not np.any(grad0.imag)
