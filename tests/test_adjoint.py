"""The adjoint (vector-Jacobian product) of the potentials, L = sum(phi^2) / N.

Closed form (sources = targets z, charges q, cotangent g = dL/dphi = 2 phi / N):
dL/dq_j = psi(z_j), dL/dz_j = q_j grad psi(z_j) + g_j grad phi(z_j), psi the potential of
the charges g.  The CPU test pins this formula against central finite differences of the
f64 direct-sum loss; the GPU tests compare the library's fmmh_vjp with the formula on
direct sums, with finite differences (h = 1e-4, the 1e-5 bar of the north star) and
through torch.autograd.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _direct(z, q):
    phi, grad = O.direct_sum(z, q, z)
    return phi / (4 * np.pi), grad / (4 * np.pi)


def _loss(z, q):
    phi, _ = _direct(z, q)
    return float(np.sum(phi * phi) / len(q))


def _formula(z, q):
    phi, field = _direct(z, q)
    g = 2.0 * phi / len(q)
    psi, dpsi = _direct(z, g)
    return psi, q[:, None] * dpsi + g[:, None] * field, g, field


def test_adjoint_formula_against_finite_differences():
    rng = O.Rng(31)
    z = rng.cube(200)
    q = rng.normals(200)
    gq, gz, _, _ = _formula(z, q)
    h = 1e-4
    for j in (0, 17, 123, 199):
        qp, qm = q.copy(), q.copy()
        qp[j] += h
        qm[j] -= h
        fd = (_loss(z, qp) - _loss(z, qm)) / (2 * h)
        assert abs(fd - gq[j]) <= 1e-7 * max(1.0, abs(gq[j])), (j, fd, gq[j])
        for k in range(3):
            zp, zm = z.copy(), z.copy()
            zp[j, k] += h
            zm[j, k] -= h
            fd = (_loss(zp, q) - _loss(zm, q)) / (2 * h)
            assert abs(fd - gz[j, k]) <= 1e-6 * max(1.0, np.abs(gz).max()), (j, k, fd, gz[j, k])


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def _gpu():
    import fmm_amd
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    return fmm_amd


@pytest.mark.gpu
@pytest.mark.parametrize("precision,p,bar", [("f64", 12, 3e-6), ("mixed", 12, 3e-6)])
def test_vjp_matches_direct_formula(_gpu, precision, p, bar):
    # the adjoint is one more FMM evaluation: its error is the FMM's at this p (the
    # gradient terms of dL/dz are the least accurate part, ~1e-6 at p = 12)
    fmm_amd = _gpu
    rng = O.Rng(32)
    z = np.vstack([rng.cube(3000), rng.plummer(1000, (0.5, 0.5, 0.5), 0.05)])
    q = rng.normals(len(z))
    f = fmm_amd.FmmBuilder(p).precision(precision).max_points_per_leaf(32).build(z)
    out = f.evaluate(q)
    g = 2.0 * out.potential / len(q)
    gq, gz = f.vjp(q, g, out.gradient)
    rq, rz, _, _ = _formula(z, q)
    eq, ez = _rel(gq, rq), _rel(gz, rz)
    print(f"vjp {precision} p={p}: rel dL/dq {eq:.3e} dL/dz {ez:.3e}", flush=True)
    assert eq < bar and ez < bar


@pytest.mark.gpu
def test_vjp_within_finite_differences(_gpu):
    # central differences of the f64 direct-sum loss, h = 1e-4, against the mixed-precision
    # FMM adjoint at p = 10: the north star's 1e-5
    fmm_amd = _gpu
    rng = O.Rng(33)
    z = rng.cube(1500)
    q = rng.normals(1500)
    f = fmm_amd.FmmBuilder(10).precision("mixed").max_points_per_leaf(32).build(z)
    out = f.evaluate(q)
    gq, gz = f.vjp(q, 2.0 * out.potential / len(q), out.gradient)
    h = 1e-4
    scale_q, scale_z = np.abs(gq).max(), np.abs(gz).max()
    worst = 0.0
    for j in (3, 400, 999, 1499):
        qp, qm = q.copy(), q.copy()
        qp[j] += h
        qm[j] -= h
        fd = (_loss(z, qp) - _loss(z, qm)) / (2 * h)
        worst = max(worst, abs(fd - gq[j]) / scale_q)
        for k in range(3):
            zp, zm = z.copy(), z.copy()
            zp[j, k] += h
            zm[j, k] -= h
            fd = (_loss(zp, q) - _loss(zm, q)) / (2 * h)
            worst = max(worst, abs(fd - gz[j, k]) / scale_z)
    print(f"vjp vs finite differences: worst {worst:.3e}", flush=True)
    assert worst <= 1e-5


@pytest.mark.gpu
def test_torch_autograd(_gpu):
    torch = pytest.importorskip("torch")
    from fmm_amd.autograd import FmmPotential, fmm_for
    rng = O.Rng(34)
    z = rng.cube(2000)
    q = rng.normals(2000)
    f = fmm_for(z, p=8, leaf_points=32)
    zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    phi = FmmPotential.apply(zt, qt, f)
    loss = (phi * phi).mean()
    loss.backward()
    out = f.evaluate(q)
    gq, gz = f.vjp(q, 2.0 * out.potential / len(q), out.gradient)
    assert zt.grad is not None and qt.grad is not None
    assert _rel(qt.grad.numpy(), gq) < 1e-12
    assert _rel(zt.grad.numpy(), gz) < 1e-12
    assert abs(loss.item() - float(np.mean(out.potential ** 2))) <= 1e-12 * loss.item()


@pytest.mark.gpu
def test_torch_autograd_device_inputs(_gpu):
    # charges and the cotangent are produced by queued torch kernels on the GPU (a long
    # matrix product first, so they are still in flight when the library is called): the
    # operator's stream must wait for torch's stream before it reads them
    torch = pytest.importorskip("torch")
    from fmm_amd.autograd import FmmPotential, fmm_for
    rng = O.Rng(35)
    z = rng.cube(3000)
    q = rng.normals(3000)
    f = fmm_for(z, p=8, leaf_points=32)
    dev = torch.device("cuda", 0)
    zt = torch.tensor(z, dtype=torch.float64, device=dev, requires_grad=True)
    base = torch.tensor(q, dtype=torch.float64, device=dev)
    big = torch.randn(2048, 2048, dtype=torch.float64, device=dev)
    for _ in range(4):
        big = big @ big / 2048.0  # keeps torch's stream busy
    qt = (base * 2.0 + 0.0 * big[0, 0]).requires_grad_(True)
    phi = FmmPotential.apply(zt, qt, f)
    w = (big.abs().sum() * 0.0 + 3.0)  # the loss weight also comes out of queued work
    loss = w * (phi * phi).mean()
    loss.backward()
    out = f.evaluate(2.0 * q)
    g = 3.0 * 2.0 * out.potential / len(q)
    gq, gz = f.vjp(2.0 * q, g, out.gradient)
    assert _rel(phi.detach().cpu().numpy(), out.potential) < 1e-12
    assert _rel(qt.grad.cpu().numpy(), gq) < 1e-12
    assert _rel(zt.grad.cpu().numpy(), gz) < 1e-12


@pytest.mark.gpu
def test_autograd_moved_points(_gpu):
    # FmmPotential refuses points other than the FMM's; Potential rebuilds and so
    # evaluates at the moved points
    torch = pytest.importorskip("torch")
    import fmm_amd
    from fmm_amd.autograd import FmmPotential, Potential, fmm_for
    rng = O.Rng(36)
    z = rng.cube(1500)
    q = rng.normals(1500)
    f = fmm_for(z, p=8, leaf_points=32)
    moved = z.copy()
    moved[7] += (0.01, -0.02, 0.005)
    zt = torch.tensor(moved, dtype=torch.float64, requires_grad=True)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    with pytest.raises(fmm_amd.FmmError):
        FmmPotential.apply(zt, qt, f)
    pot = Potential(p=8, leaf_points=32)
    phi0 = pot(torch.tensor(z, dtype=torch.float64), qt).detach().cpu().numpy()
    phi1 = pot(zt, qt)
    assert pot.builds == 2
    ref, _ = _direct(moved, q)
    assert _rel(phi1.detach().cpu().numpy(), ref) < 1e-5
    assert np.abs(phi1.detach().cpu().numpy() - phi0).max() > 1e-4
    (phi1 * phi1).mean().backward()
    assert zt.grad is not None and np.all(np.isfinite(zt.grad.numpy()))


@pytest.mark.gpu
def test_device_buffers_checked(_gpu):
    # device pointers of the wrong length, dtype or layout are refused before any launch
    torch = pytest.importorskip("torch")
    import fmm_amd
    rng = O.Rng(37)
    z = rng.cube(500)
    f = fmm_amd.FmmBuilder(4).max_points_per_leaf(32).build(z)
    op = f.operator
    dev = torch.device("cuda", 0)
    for bad in (torch.zeros(499, dtype=torch.float64, device=dev), torch.zeros(500, dtype=torch.float32, device=dev),
                torch.zeros(1000, dtype=torch.float64, device=dev)[::2]):
        with pytest.raises(fmm_amd.FmmError):
            op.load_charges(bad)
    op.load_charges(torch.ones(500, dtype=torch.float64, device=dev))
    op.reset()
    op.evaluate()
    with pytest.raises(fmm_amd.FmmError):
        op.read_output_device(torch.zeros(500, dtype=torch.float64, device=dev),
                              torch.zeros(499, 3, dtype=torch.float64, device=dev))
    phi = torch.zeros(500, dtype=torch.float64, device=dev)
    op.read_output_device(phi)
    ref, _ = _direct(z, np.ones(500))
    assert _rel(phi.cpu().numpy(), ref) < 1e-4
