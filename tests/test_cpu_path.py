"""The CPU baseline: the reference's multi-threaded host evaluation restated in C
(oracle/cpu_path.c: threads over targets, targets-in-lanes AVX-512 P2P with vrsqrt14 and
two Newton steps, the batched M2L column kernel with fused steps).

Pinned against the serial oracle on the same tree (the fused M2L chain and the Newton
inverse square root differ from it only in rounding: 1e-13 relative), for one and for
several threads (each target's body is serial, so the thread count changes no bit), on a
uniform and on an adaptive tree with W/X lists and the near rule, and against the
reference's published errors of configuration 1 (N = 100,000, p = 6, 64 per leaf:
1.42e-4 / 1.30e-4 to three digits, fmm-bench/results/phase4s-gh200.md).
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O


def _rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("case", ["cube", "adaptive"])
def test_cpu_path_matches_serial_oracle(case):
    rng = O.Rng(61)
    if case == "cube":
        pts = rng.cube(20000)
        f = O.Fmm(pts, p=8, leaf_points=40)
    else:
        pts = np.vstack([rng.plummer(6000, (0.5, 0.5, 0.5), 0.05), rng.sphere(4000, (0.5, 0.5, 0.5), 0.5)])
        f = O.Fmm(pts, p=7, leaf_points=24, near_rule=(32.0, 32.0))
    q = rng.normals(len(pts))
    phi, grad = f.evaluate(q)
    one = f.evaluate_cpu_path(q, threads=1)
    many = f.evaluate_cpu_path(q, threads=4)
    assert np.array_equal(one[0], many[0]) and np.array_equal(one[1], many[1])
    assert _rel_max(one[0], phi) < 1e-13
    assert _rel_max(one[1], grad) < 1e-13


def test_cpu_path_p2p_known_answer():
    # fmm-ref/src/p2p.rs:51-60 through a whole evaluation: two points in one leaf, where
    # the FMM is exactly the direct sum (phi = (-1/5, 2/5) / 4 pi)
    pts = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    q = np.array([2.0, -1.0])
    f = O.Fmm(pts, p=3, leaf_points=64)
    phi, grad = f.evaluate_cpu_path(q, threads=2)
    assert np.allclose(phi * 4 * np.pi, [-1.0 / 5.0, 2.0 / 5.0], rtol=1e-15, atol=0)
    assert np.allclose(grad[1] * 4 * np.pi, [-6.0 / 125.0, -8.0 / 125.0, 0.0], rtol=1e-14, atol=1e-17)


def test_cpu_path_reproduces_published_errors(golden_dir):
    pub = json.load(open(os.path.join(golden_dir, "published.json")))
    row = [r for r in pub["errors"] if r["n"] == 100000 and r["p"] == 6][0]
    pts, q, sample = O.bench_problem(100000, 1000)
    f = O.Fmm(pts, p=6, leaf_points=64)
    phi, grad = f.evaluate_cpu_path(q)
    ephi, egrad = O.direct_sum(pts, q, pts[sample])
    e_phi = O.error_norms(phi[sample], ephi / (4 * np.pi))[0]
    e_grad = O.error_norms(grad[sample], egrad / (4 * np.pi), True)[0]
    assert abs(e_phi - row["phi"]) <= 0.005 * row["phi"], e_phi
    assert abs(e_grad - row["grad"]) <= 0.005 * row["grad"], e_grad
