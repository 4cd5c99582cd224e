"""Batched independent clouds (the 256-cloud Gaussian-blob workload) against the oracle.

A forest of Gaussian-blob clouds of log-uniform sizes (the generator of the benchmark's
batched workload, fmm_amd.points.batched_clouds; blobs as fmm-validate/src/points.rs:114-135)
at p = 6, at most 64 points per leaf, the near rule at (p+1)^2 / 2:

- each cloud's slice of the forest's output against the oracle FMM of that cloud alone
  (same tree rules, lists and near rule): f64 to 1e-12, mixed to 1e-6;
- the forest's adjoint of L = sum(phi^2) / N against the oracle's adjoint on the same
  trees (the oracle FMM of the charges g = 2 phi / N, combined by the closed form
  dL/dq_j = psi(z_j), dL/dz_j = q_j grad psi(z_j) + g_j grad phi(z_j)): f64 to 1e-12;
- the same adjoint against the closed form on the oracle's f64 direct sums (the FMM's
  truncation at p = 6 bounds the difference) and against central finite differences of
  the direct-sum loss, h = 1e-4;
- the full 256-cloud forest (14.25M points): finite outputs, and the error against the
  compensated direct sum at sampled targets of every cloud within the p = 6 FMM error.
"""
import numpy as np
import pytest

import fmm_amd
from fmm_amd import points as P
from oracle import oracle as O

p = 6
LEAF = 64
NEAR = (0.5 * (p + 1) ** 2,) * 2


@pytest.fixture(scope="module")
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def small_forest():
    # 8 clouds, sizes log-uniform in [2000, 16000]: the oracle finishes each in a second
    clouds, charges = P.batched_clouds(nclouds=8, nmin=2000, nmax=16000, seed=P.BENCH_SEED + 11)
    first = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    oracles = [O.Fmm(c, p=p, leaf_points=LEAF, near_rule=NEAR) for c in clouds]
    ref = [o.evaluate(q) for o, q in zip(oracles, charges)]
    return clouds, charges, first, oracles, ref


def _forest(clouds, precision):
    b = fmm_amd.FmmBuilder(p).precision(precision).max_points_per_leaf(LEAF).near_rule("device")
    return b.build_batched(clouds)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bar", [("f64", 1e-12), ("mixed", 1e-6)])
def test_forest_against_oracle_per_cloud(_gpu, small_forest, precision, bar):
    clouds, charges, first, oracles, ref = small_forest
    forest = _forest(clouds, precision)
    out = forest.evaluate(np.concatenate(charges))
    assert np.all(np.isfinite(out.potential)) and np.all(np.isfinite(out.gradient))
    for c, (rphi, rgrad) in enumerate(ref):
        sl = slice(first[c], first[c + 1])
        e1, e2 = _rel(out.potential[sl], rphi), _rel(out.gradient[sl], rgrad)
        print(f"forest {precision} cloud {c} (n={len(clouds[c])}): vs oracle phi {e1:.2e} grad {e2:.2e}",
              flush=True)
        assert e1 < bar and e2 < bar


def _oracle_adjoint(o, z, q, phi, field, g):
    """dL/dq and dL/dz of the oracle FMM on the same tree (the closed form with the
    oracle's fields)."""
    psi, dpsi = o.evaluate(g)
    return psi, q[:, None] * dpsi + g[:, None] * field


def _direct(z, q):
    phi, grad = O.direct_sum(z, q, z)
    return phi / (4 * np.pi), grad / (4 * np.pi)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_forest_adjoint(_gpu, small_forest, precision):
    clouds, charges, first, oracles, ref = small_forest
    forest = _forest(clouds, precision)
    q = np.concatenate(charges)
    out = forest.evaluate(q)
    ntot = len(q)
    g = 2.0 * out.potential / ntot
    gq, gz = forest.vjp(q, g, out.gradient)
    worst_oracle = worst_closed = 0.0
    smallest = set(np.argsort([len(z) for z in clouds])[:3].tolist())
    for c, (o, z) in enumerate(zip(oracles, clouds)):
        sl = slice(first[c], first[c + 1])
        rphi, rgrad = ref[c]
        # the oracle's adjoint on the same tree, fed the oracle's own forward output
        oq, oz = _oracle_adjoint(o, z, charges[c], rphi, rgrad, 2.0 * rphi / ntot)
        worst_oracle = max(worst_oracle, _rel(gq[sl], oq), _rel(gz[sl], oz))
        if c in smallest:
            # the closed form on f64 direct sums: the FMM's truncation at p = 6
            dphi, dfield = _direct(z, charges[c])
            dg = 2.0 * dphi / ntot
            psi, dpsi = _direct(z, dg)
            cq, cz = psi, charges[c][:, None] * dpsi + dg[:, None] * dfield
            worst_closed = max(worst_closed, _rel(gq[sl], cq), _rel(gz[sl], cz))
    print(f"forest adjoint {precision}: vs oracle adjoint {worst_oracle:.2e}, vs direct-sum closed form "
          f"{worst_closed:.2e}", flush=True)
    assert worst_oracle < (1e-12 if precision == "f64" else 2e-6)
    assert worst_closed < 1e-3  # p = 6: the FMM's own error (published 1.4e-4 on the cube)


@pytest.mark.gpu
@pytest.mark.parametrize("degree,bar", [(6, 5e-5), (10, 1e-5)])
def test_forest_adjoint_finite_differences(_gpu, small_forest, degree, bar):
    # central differences (h = 1e-4) of the f64 direct-sum loss L = sum(phi^2) / N over the
    # whole forest; the clouds do not interact, so a coordinate of cloud c moves only
    # cloud c's terms, and the FD of L is that of cloud c's part (N the forest's count).
    # The adjoint is one more FMM evaluation, so it carries the FMM's truncation: at the
    # batched workload's p = 6 (relative L2 error ~1e-4) the worst component differs from
    # the FD by a few 1e-5 of the largest; at p = 10 by less than 1e-5
    clouds, charges, first, oracles, ref = small_forest
    b = fmm_amd.FmmBuilder(degree).precision("mixed").max_points_per_leaf(LEAF).near_rule("device")
    forest = b.build_batched(clouds)
    q = np.concatenate(charges)
    ntot = len(q)
    out = forest.evaluate(q)
    gq, gz = forest.vjp(q, 2.0 * out.potential / ntot, out.gradient)
    c = int(np.argmin([len(z) for z in clouds]))
    z, qc = clouds[c], charges[c]
    sl = slice(first[c], first[c + 1])
    gqc, gzc = gq[sl], gz[sl]

    def loss(zz, qq):
        phi, _ = _direct(zz, qq)
        return float(np.sum(phi * phi) / ntot)

    h = 1e-4
    scale_q, scale_z = np.abs(gqc).max(), np.abs(gzc).max()
    worst = 0.0
    for j in P.SplitMix64(77).sample(len(z), 6):
        qp, qm = qc.copy(), qc.copy()
        qp[j] += h
        qm[j] -= h
        fd = (loss(z, qp) - loss(z, qm)) / (2 * h)
        worst = max(worst, abs(fd - gqc[j]) / scale_q)
        for k in range(3):
            zp, zm = z.copy(), z.copy()
            zp[j, k] += h
            zm[j, k] -= h
            fd = (loss(zp, qc) - loss(zm, qc)) / (2 * h)
            worst = max(worst, abs(fd - gzc[j, k]) / scale_z)
    print(f"forest adjoint vs finite differences (cloud {c}, n={len(z)}, p = {degree}): worst {worst:.3e}",
          flush=True)
    assert worst <= bar


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_full_batched_workload(_gpu, precision):
    # the benchmark's 256 clouds (2k..256k points each, 14.25M in all)
    clouds, charges = P.batched_clouds()
    n = int(sum(len(c) for c in clouds))
    forest = _forest(clouds, precision)
    out = forest.evaluate(np.concatenate(charges))
    assert len(out.potential) == n
    assert np.all(np.isfinite(out.potential)) and np.all(np.isfinite(out.gradient))
    first = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    num = den = numg = deng = 0.0
    for c, (pts, q) in enumerate(zip(clouds, charges)):
        idx = P.SplitMix64(1000 + c).sample(len(pts), 8)
        e, eg = fmm_amd.direct_sum(pts, q, pts[idx])
        e, eg = e / (4 * np.pi), eg / (4 * np.pi)
        num += float(np.sum((out.potential[first[c] + idx] - e) ** 2))
        den += float(np.sum(e ** 2))
        numg += float(np.sum((out.gradient[first[c] + idx] - eg) ** 2))
        deng += float(np.sum(eg ** 2))
    e_phi, e_grad = (num / den) ** 0.5, (numg / deng) ** 0.5
    print(f"256-cloud forest {precision} (n={n}): error vs direct sum phi {e_phi:.2e} grad {e_grad:.2e}", flush=True)
    assert e_phi < 5e-4 and e_grad < 5e-4
