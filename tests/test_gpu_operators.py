"""Operator-by-operator GPU parity: multipoles and locals of every level after an
evaluation, and each leaf kind alone, against the oracle (f64, same tree and lists).

The expansions are compared through fmmh_read_expansions; each level's relative error
isolates P2M/M2M (multipoles) and L2L/M2L/P2L (locals).  Leaf kinds run alone through
the level calls of the C ABI after a reset.
"""
import sys

import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def _rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    s = np.linalg.norm(np.asarray(b))
    return d / s if s > 0 else d


@pytest.fixture(scope="module")
def setup():
    rng = O.Rng(1234)
    pts = np.vstack([rng.cube(4000), rng.plummer(3000, (0.5, 0.5, 0.5), 0.06)])
    q = rng.normals(len(pts))
    p, leaf = 6, 24
    fmm = fmm_amd.FmmBuilder(p).max_points_per_leaf(leaf).build(pts)
    ref = O.Fmm(pts, p=p, leaf_points=leaf)
    ref.evaluate(q)
    fmm.evaluate(q)
    return pts, q, p, leaf, fmm, ref


def test_multipoles_every_level(setup):
    pts, q, p, leaf, fmm, ref = setup
    for level in range(fmm.plan.nlevels - 1, -1, -1):
        a = fmm.operator.expansions(False, level)
        b = ref.expansions(False, level)
        e = _rel(a, b)
        _log(f"multipoles level {level}: boxes {len(a)} rel {e:.3e}")
        assert e < 1e-13, level


def test_locals_every_level(setup):
    pts, q, p, leaf, fmm, ref = setup
    for level in range(fmm.plan.nlevels):
        a = fmm.operator.expansions(True, level)
        b = ref.expansions(True, level)
        e = _rel(a, b)
        _log(f"locals level {level}: boxes {len(a)} rel {e:.3e}")
        assert e < 1e-13, level


def _near_field(plan, pts, q, fmm):
    """Direct sums over each target leaf's near row (coincident points skipped), /4pi."""
    leaves = plan.point_leaves()
    lk = plan.leaf_keys()
    levels = (lk & np.uint64(0x7FFF)).astype(int)
    members = {}
    for i, lf in enumerate(leaves):
        members.setdefault(int(lf), []).append(i)
    phi = np.zeros(len(pts))
    grad = np.zeros((len(pts), 3))
    offsets = np.concatenate([[0], np.cumsum(np.bincount(levels, minlength=plan.nlevels))])
    for level in range(plan.nlevels):
        off, ent = plan.list("near", level)
        for r in range(len(off) - 1):
            lf = offsets[level] + r
            t = members.get(int(lf), [])
            if not t:
                continue
            s = [j for e in ent[off[r]:off[r + 1]] for j in members.get(int(e), [])]
            if not s:
                continue
            f, g = O.direct_sum(pts[s], q[s], pts[t])
            phi[t] += f / (4 * np.pi)
            grad[t] += g / (4 * np.pi)
    return phi, grad


def test_p2p_alone(setup):
    pts, q, p, leaf, fmm, ref = setup
    op = fmm.operator
    op.load_charges(q)
    op.reset()
    for level in range(fmm.plan.nlevels):
        op.call("p2p", level)
    out = op.read_output()
    phi, grad = _near_field(fmm.plan, pts, q, fmm)
    e1, e2 = _rel(out.potential, phi), _rel(out.gradient, grad)
    _log(f"p2p alone: rel phi {e1:.3e} grad {e2:.3e}")
    assert e1 < 1e-13 and e2 < 1e-12


def test_leaf_kinds_alone(setup):
    # L2P and M2P alone after a full upward/downward pass: the difference of the full
    # output and P2P alone is L2P + M2P; compare it with the oracle's
    pts, q, p, leaf, fmm, ref = setup
    op = fmm.operator
    full = fmm.evaluate(q)
    op.load_charges(q)
    op.reset()
    for level in range(fmm.plan.nlevels):
        op.call("p2p", level)
    near = op.read_output()
    rphi, rgrad = ref.evaluate(q)
    nphi, ngrad = _near_field(fmm.plan, pts, q, fmm)
    e1 = _rel(full.potential - near.potential, rphi - nphi)
    e2 = _rel(full.gradient - near.gradient, rgrad - ngrad)
    _log(f"far field (l2p+m2p): rel phi {e1:.3e} grad {e2:.3e}")
    # the far field is formed here as a difference of two totals whose rounding (~1e-14
    # of the near field) is measured against the smaller far part: 1e-11 is that bar
    assert e1 < 1e-11 and e2 < 1e-11


@pytest.fixture(scope="module")
def mixed_pair():
    # p = 12 (the 1e-6 setting) on a cube plus a cluster: dense levels run as tiles,
    # the refined cluster levels as pairs into f32 slots
    rng = O.Rng(77)
    pts = np.vstack([rng.cube(20000), rng.plummer(10000, (0.3, 0.6, 0.5), 0.04)])
    q = rng.normals(len(pts))
    plain = fmm_amd.FmmBuilder(12).max_points_per_leaf(32).build(pts)
    mixed = fmm_amd.FmmBuilder(12).precision("mixed").max_points_per_leaf(32).build(pts)
    return pts, q, plain, mixed


def test_mixed_m2l_within_split_precision(mixed_pair):
    # M2L alone in bf16x3 (six split products, f32 per offset, f64 across offsets)
    # against M2L in f64 on the same multipoles: the split keeps 24 bits of each factor,
    # so the locals agree to the f32 contract of 1e-6 (laplace-fmm-plan.md:1696)
    pts, q, plain, mixed = mixed_pair
    L = plain.plan.nlevels - 1
    st = mixed.operator.stats()
    _log(f"  m2l pairs {st['m2l_pairs']} slots {st['m2l_slots']} levels {L + 1}")
    locs = []
    for f in (plain, mixed):
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
        _log(f"mixed m2l level {level}: rel {e:.3e}")
        assert e < 1e-6, level


def test_mixed_p2p_within_f32(mixed_pair):
    # P2P alone in f32 with double-float separations and f64 tile sums, against the
    # f64 near-field direct sums
    pts, q, plain, mixed = mixed_pair
    op = mixed.operator
    op.load_charges(q)
    op.reset()
    for level in range(mixed.plan.nlevels):
        op.call("p2p", level)
    out = op.read_output()
    phi, grad = _near_field(mixed.plan, pts, q, mixed)
    e1, e2 = _rel(out.potential, phi), _rel(out.gradient, grad)
    _log(f"mixed p2p alone: rel phi {e1:.3e} grad {e2:.3e}")
    assert e1 < 1e-6 and e2 < 1e-6


@pytest.mark.parametrize("leaf", [48, 128])
def test_mixed_p2p_uniform_chunks(leaf):
    # Leaves of more than 32 targets take the mixed P2P's uniform-source path (one
    # target per lane up to 64, two above); coincident copies and pairs 1e-9 apart
    # check the r^2 == 0 select and the double-float separations against f64
    _log(f"case mixed p2p uniform chunks leaf={leaf}")
    rng = O.Rng(4242)
    base = rng.cube(16000 if leaf <= 64 else 50000)
    near = base[:300] + 1e-9 * rng.cube(300)
    pts = np.vstack([base, base[300:500], near])
    q = rng.normals(len(pts))
    mixed = fmm_amd.FmmBuilder(8).precision("mixed").max_points_per_leaf(leaf).build(pts)
    sizes = np.bincount(mixed.plan.point_leaves())
    assert sizes.max() > (64 if leaf > 64 else 32)
    op = mixed.operator
    op.load_charges(q)
    op.reset()
    for level in range(mixed.plan.nlevels):
        op.call("p2p", level)
    out = op.read_output()
    phi, grad = _near_field(mixed.plan, pts, q, mixed)
    assert np.all(np.isfinite(out.potential)) and np.all(np.isfinite(out.gradient))
    e1, e2 = _rel(out.potential, phi), _rel(out.gradient, grad)
    _log(f"  largest leaf {sizes.max()}: rel phi {e1:.3e} grad {e2:.3e}")
    assert e1 < 1e-6 and e2 < 1e-6


def _frame(s_key, t_key):
    """c(s|t), r(s|t): the source box's centre and radius in the target leaf's units, exact
    from the integer cells (CONVENTIONS §3.13)."""
    ls, *cs = O.decode(int(s_key))
    lt, *ct = O.decode(int(t_key))
    L = max(ls, lt)
    c = [((2 * cs[k] + 1) * 2 ** (L - ls) - (2 * ct[k] + 1) * 2 ** (L - lt)) * 2.0 ** (lt - L) for k in range(3)]
    return np.array(c), 2.0 ** (lt - ls)


def _kind_alone_reference(ref, kind):
    """L2P alone (each leaf's local at its targets) or M2P alone (each leaf's W-list
    multipoles at its targets) from the oracle's own expansions and leaf operators,
    divided by 4 pi r_t (r_t^2 for the gradient), in caller order."""
    p = ref.p
    lk = ref.leaf_keys()
    levels = (lk & np.uint64(0x7FFF)).astype(int)
    u, order = ref.leaf_coordinates(targets=True)
    pl = ref.point_leaves(targets=True)
    start = np.concatenate([[0], np.cumsum(np.bincount(pl, minlength=len(lk)))])
    box, side = ref.domain()
    phi = np.zeros(len(order))
    grad = np.zeros((len(order), 3))
    offsets = np.concatenate([[0], np.cumsum(np.bincount(levels, minlength=ref.nlevels))])
    keys = [ref.box_keys(l) for l in range(ref.nlevels)]
    for l in range(ref.nlevels):
        rt = side / 2.0 ** (l + 1)
        if kind == "l2p":
            loc = ref.expansions(True, l)
        else:
            woff, went, wlev = ref.list("w", l, aux=True)  # W entries: box index and its level
        for r in range(offsets[l + 1] - offsets[l]):
            lf = offsets[l] + r
            t0, t1 = start[lf], start[lf + 1]
            if t0 == t1:
                continue
            tgt = u[t0:t1]
            f = np.zeros(t1 - t0)
            g = np.zeros((t1 - t0, 3))
            if kind == "l2p":
                b = int(np.searchsorted(keys[l], lk[lf]))
                f, g = O.l2p(p, (0.0, 0.0, 0.0), 1.0, loc[b], tgt)
            else:
                for e in range(woff[r], woff[r + 1]):
                    sb = int(went[e])
                    sl = int(wlev[e])
                    c, rr = _frame(keys[sl][sb], lk[lf])
                    df, dg = O.m2p(p, c, rr, ref.expansions(False, sl)[sb], tgt)
                    f = f + df
                    g = g + dg
            phi[order[t0:t1]] = f / (4 * np.pi * rt)
            grad[order[t0:t1]] = g / (4 * np.pi * rt * rt)
    return phi, grad


@pytest.mark.parametrize("kind", ["l2p", "m2p"])
def test_leaf_kind_alone_against_oracle_operators(setup, kind):
    # the upward and downward passes through the level calls (multipoles and locals equal
    # the oracle's to 1e-13), then only the L2P or only the M2P level calls: the output is
    # that kind alone, compared with the oracle's leaf operator on the oracle's expansions
    pts, q, p, leaf, fmm, ref = setup
    op = fmm.operator
    L = fmm.plan.nlevels - 1
    op.load_charges(q)
    op.reset()
    for level in range(L, -1, -1):
        op.call("p2m", level)
        if level >= 1:
            op.call("m2m", level - 1)
    for level in range(1, L + 1):
        for k in ("l2l", "m2l", "p2l"):
            op.call(k, level)
    for level in range(L + 1):
        op.call(kind, level)
    out = op.read_output()
    rphi, rgrad = _kind_alone_reference(ref, kind)
    assert np.linalg.norm(rphi) > 0, "the tree must exercise this kind"
    e1, e2 = _rel(out.potential, rphi), _rel(out.gradient, rgrad)
    _log(f"{kind} alone vs oracle operators: rel phi {e1:.3e} grad {e2:.3e}")
    assert e1 < 1e-13 and e2 < 1e-13
