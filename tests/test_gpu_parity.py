"""GPU parity: the HIP path through the C ABI against the oracle (f64 restatement of the
reference FMM) and the f64 direct sum.

Bars: the plain f64 GPU path within 1e-12 relative L2 of the oracle FMM on the same
tree; the mixed path (f32 P2P) within its precision's contract of the oracle; the
reference's published errors reproduced against the direct sum.
"""
import json
import os
import sys

import numpy as np
import pytest

import fmm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b, vec=False):
    return O.error_norms(a, b, vec)[0]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


CASES = [
    ("cube", 6000, 3, 40),
    ("cube", 8000, 6, 48),
    ("plummer", 8000, 8, 32),
    ("mixed", 10000, 12, 64),
    ("cube", 4000, 15, 40),
]


def _points(kind, n, seed=11):
    rng = O.Rng(seed)
    if kind == "cube":
        return rng.cube(n)
    if kind == "plummer":
        return rng.plummer(n, (0.5, 0.5, 0.5), 0.05)
    return np.vstack([rng.sphere(n // 2, (0.5, 0.5, 0.5), 0.5), rng.plummer(n - n // 2, (0.5, 0.5, 0.5), 0.1)])


@pytest.mark.parametrize("kind,n,p,leaf", CASES)
def test_plain_f64_matches_oracle(kind, n, p, leaf):
    _log(f"case plain f64 {kind} n={n} p={p} leaf={leaf}")
    pts = _points(kind, n)
    q = O.Rng(5).normals(n)
    fmm = fmm_amd.FmmBuilder(p).max_points_per_leaf(leaf).build(pts)
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=p, leaf_points=leaf)
    rphi, rgrad = ref.evaluate(q)
    assert _rel(out.potential, rphi) < 1e-12
    assert _rel(out.gradient, rgrad, True) < 1e-12
    # deterministic: a second evaluation gives the same bits
    out2 = fmm.evaluate(q)
    assert np.array_equal(out.potential, out2.potential)
    assert np.array_equal(out.gradient, out2.gradient)


@pytest.mark.parametrize("rule", [(float("inf"), float("inf")), (20.0, 10.0)])
def test_near_rule_plain_f64_matches_oracle(rule):
    _log(f"case near rule {rule}")
    rng = O.Rng(41)
    pts = np.vstack([rng.plummer(5000, (0.5, 0.5, 0.5), 0.05), rng.cube(3000)])
    q = rng.normals(len(pts))
    fmm = fmm_amd.FmmBuilder(7).max_points_per_leaf(24).near_rule(rule).build(pts)
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=7, leaf_points=24, near_rule=rule)
    rphi, rgrad = ref.evaluate(q)
    assert _rel(out.potential, rphi) < 1e-12
    assert _rel(out.gradient, rgrad, True) < 1e-12


def test_level_calls_equal_evaluate():
    _log("case level calls")
    pts = _points("plummer", 6000, 3)
    q = O.Rng(6).normals(len(pts))
    fmm = fmm_amd.FmmBuilder(6).max_points_per_leaf(24).build(pts)
    a = fmm.evaluate(q)
    b = fmm.evaluate_levels(q)
    # evaluate() adds each local's M2L before its L2L (one launch for every level's
    # M2L); the level calls add them in the reference's order: equal up to rounding
    assert _rel(a.potential, b.potential) < 1e-14
    assert _rel(a.gradient, b.gradient, True) < 1e-14
    c = fmm.evaluate_levels(q)
    assert np.array_equal(b.potential, c.potential)


def test_separate_targets_and_no_gradients():
    _log("case targets != sources, no gradients")
    rng = O.Rng(8)
    src = rng.plummer(5000, (0.5, 0.5, 0.5), 0.08)
    tgt = rng.cube(3000)
    q = rng.charges(len(src))
    fmm = fmm_amd.FmmBuilder(7).gradients(False).max_points_per_leaf(30).build(src, tgt)
    out = fmm.evaluate(q)
    assert out.gradient is None
    ref = O.Fmm(src, tgt, p=7, leaf_points=30, gradients=False)
    rphi, _ = ref.evaluate(q)
    assert _rel(out.potential, rphi) < 1e-12


def test_coincident_points_and_edge_cases():
    _log("case duplicates, zero charges, root leaf")
    rng = O.Rng(9)
    pts = rng.cube(3000)
    pts = np.vstack([pts, pts[:700]])  # coincident pairs are skipped
    q = rng.normals(len(pts))
    q[::5] = 0.0
    fmm = fmm_amd.FmmBuilder(5).max_points_per_leaf(16).build(pts)
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=5, leaf_points=16)
    rphi, rgrad = ref.evaluate(q)
    assert np.all(np.isfinite(out.potential)) and np.all(np.isfinite(out.gradient))
    assert _rel(out.potential, rphi) < 1e-12
    assert _rel(out.gradient, rgrad, True) < 1e-12
    # a tree that is only the root: P2P alone, against the direct sum
    few = rng.cube(40)
    qf = rng.charges(40)
    f2 = fmm_amd.FmmBuilder(4).max_points_per_leaf(64).build(few)
    assert f2.plan.nlevels == 1
    o2 = f2.evaluate(qf)
    e, g = O.direct_sum(few, qf, few)
    assert _rel(o2.potential, e / (4 * np.pi)) < 1e-14
    assert _rel(o2.gradient, g / (4 * np.pi), True) < 1e-13
    # all charges zero
    o3 = fmm.evaluate(np.zeros(len(pts)))
    assert not np.any(o3.potential) and not np.any(o3.gradient)


def test_mixed_precision_within_f32_contract():
    _log("case mixed precision")
    pts = _points("cube", 8000, 21)
    q = O.Rng(22).normals(len(pts))
    fmm = fmm_amd.FmmBuilder(8).precision("mixed").max_points_per_leaf(48).build(pts)
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=8, leaf_points=48)
    rphi, rgrad = ref.evaluate(q)
    # f32 P2P with double-float separations: the reference's contract for f32 sums is
    # 1e-6 relative to the f64 direct sum (docs/design/laplace-fmm-plan.md:1696)
    assert _rel(out.potential, rphi) < 1e-6
    assert _rel(out.gradient, rgrad, True) < 1e-6


def test_mixed_precision_adaptive_tree():
    _log("case mixed precision adaptive")
    pts = _points("mixed", 12000, 23)
    q = O.Rng(24).normals(len(pts))
    fmm = fmm_amd.FmmBuilder(10).precision("mixed").max_points_per_leaf(40).build(pts)
    st = fmm.operator.stats()
    out = fmm.evaluate(q)
    ref = O.Fmm(pts, p=10, leaf_points=40)
    rphi, rgrad = ref.evaluate(q)
    _log(f"  m2l pairs {st['m2l_pairs']} slots {st['m2l_slots']}")
    assert _rel(out.potential, rphi) < 1e-6
    assert _rel(out.gradient, rgrad, True) < 1e-6


def test_published_error_reproduced_on_gpu():
    _log("case published N=1e5 p=6")
    pub = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))
    row = [r for r in pub["errors"] if r["n"] == 100000 and r["p"] == 6][0]
    pts, q, sample = O.bench_problem(100000, 1000)
    fmm = fmm_amd.FmmBuilder(6).max_points_per_leaf(64).build(pts)
    out = fmm.evaluate(q)
    ephi, egrad = O.direct_sum(pts, q, pts[sample])
    e_phi = _rel(out.potential[sample], ephi / (4 * np.pi))
    e_grad = _rel(out.gradient[sample], egrad / (4 * np.pi), True)
    assert abs(e_phi - row["phi"]) <= 0.005 * row["phi"], e_phi
    assert abs(e_grad - row["grad"]) <= 0.005 * row["grad"], e_grad


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_graph_replays_equal_eager(precision):
    # reset + evaluate captured once in a hipGraph and replayed, with eager evaluations
    # and synchronisations in between: every replay gives the eager bits
    _log(f"case graph replays {precision}")
    pts = _points("mixed", 9000, 61)
    q = O.Rng(62).normals(len(pts))
    fmm = fmm_amd.FmmBuilder(8).precision(precision).max_points_per_leaf(32).build(pts)
    eager = fmm.evaluate(q)
    op = fmm.operator
    q2 = O.Rng(63).normals(len(pts))
    for rep in range(3):
        op.load_charges(q)
        op.evaluate_graph()
        out = op.read_output()
        assert np.array_equal(out.potential, eager.potential), rep
        assert np.array_equal(out.gradient, eager.gradient), rep
        other = fmm.evaluate(q2)  # eager work between replays
        assert np.all(np.isfinite(other.potential))


@pytest.mark.parametrize("precision,bar", [("f64", 1e-12), ("mixed", 1e-6)])
def test_graph_replay_matches_oracle(precision, bar):
    # the replayed graph's output against the oracle FMM on the same tree directly, with
    # new charges loaded before each replay (the graph reads the charges buffer, it does
    # not capture the values)
    _log(f"case graph replay vs oracle {precision}")
    pts = _points("mixed", 9000, 64)
    fmm = fmm_amd.FmmBuilder(8).precision(precision).max_points_per_leaf(32).build(pts)
    ref = O.Fmm(pts, p=8, leaf_points=32)
    op = fmm.operator
    for seed in (65, 66):
        q = O.Rng(seed).normals(len(pts))
        op.load_charges(q)
        op.evaluate_graph()
        out = op.read_output()
        rphi, rgrad = ref.evaluate(q)
        e1, e2 = _rel(out.potential, rphi), _rel(out.gradient, rgrad, True)
        _log(f"  graph {precision} seed {seed}: vs oracle phi {e1:.2e} grad {e2:.2e}")
        assert e1 < bar and e2 < bar
