"""The kernel-level entry points (the reference's per-level launch wrappers,
fmm-kernels/src/p2p.rs:1244, leaf.rs:1851-1951, movement.rs:411,482): each called through
the C ABI on device arrays laid out from a plan, against the operator's own level calls on
the same plan (the same bits), and the P2P also against the oracle's direct sum over each
leaf's near row."""
import ctypes as C

import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to("cuda:0")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _csr(plan, kind, levels, base_of_entry=None):
    """Concatenated rows of `kind` over `levels` (offsets shifted, entries mapped)."""
    off, ent = [0], []
    for l in levels:
        o, e = plan.list(kind, l)[:2]
        e = e if base_of_entry is None else base_of_entry(l, e)
        off.extend((o[1:] + len(ent)).tolist())
        ent.extend(np.asarray(e).tolist())
    return np.array(off, np.uint32), np.array(ent if ent else [0], np.uint32)


@pytest.fixture(scope="module")
def case():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    rng = O.Rng(21)
    n, p, leaf = 4000, 6, 24
    pts = np.vstack([rng.plummer(n // 2, (0.3, 0.6, 0.5), 0.05), rng.cube(n - n // 2)])
    q = O.Rng(4).normals(n)
    fmm = fmm_amd.FmmBuilder(p).max_points_per_leaf(leaf).build(pts)
    plan, op = fmm.plan, fmm.operator
    L = plan.nlevels
    nb = [len(plan.box_keys(l)) for l in range(L)]
    box_base = np.concatenate([[0], np.cumsum(nb)]).astype(np.int64)
    lkeys = plan.leaf_keys()
    cells = plan.leaf_cells()
    leaf_level = cells[:, 0].astype(np.int64)
    leaf_box = np.array([box_base[lv] + np.searchsorted(plan.box_keys(lv), k) for lv, k in zip(leaf_level, lkeys)])
    box_cells = np.vstack([plan.box_cells(l) for l in range(L)])
    us, sorder = plan.leaf_coordinates(False)
    ut, torder = plan.leaf_coordinates(True)
    d = dict(pts=pts, q=q, p=p, plan=plan, op=op, L=L, box_base=box_base, cells=cells, leaf_box=leaf_box,
             box_cells=box_cells, us=us, ut=ut, sorder=sorder, torder=torder,
             sstart=plan.point_starts(False), tstart=plan.point_starts(True), nc=(p + 1) ** 2)
    _, side = plan.domain()
    d["leaf_r"] = side / 2.0 ** (leaf_level + 1)
    return d


def _run_operator(c, kinds_by_level, read=None):
    op = c["op"]
    op.load_charges(c["q"])
    op.reset()
    for kind, levels in kinds_by_level:
        for l in levels:
            op.call(kind, l)
    if read == "output":
        return op.read_output()
    return np.vstack([op.expansions(read == "locals", l) for l in range(c["L"])])


def _gather(c, phi_hat, grad_hat):
    import torch

    n = len(c["pts"])
    phi = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    grad = torch.zeros(3 * n, dtype=torch.float64, device="cuda:0")
    tstart, leaf_r, order = _dev(c["tstart"]), _dev(c["leaf_r"]), _dev(c["torder"])
    _lib.check(_lib.lib().fmmh_kernel_gather_output(0, len(c["cells"]), _ptr(tstart), _ptr(leaf_r), _ptr(order),
                                                    _ptr(phi_hat), _ptr(grad_hat), 1, _ptr(phi), _ptr(grad), None))
    torch.cuda.synchronize()
    return phi.cpu().numpy(), grad.cpu().numpy().reshape(-1, 3)


def _points(c):
    s, t = c["us"], c["ut"]
    return [_dev(s[:, 0]), _dev(s[:, 1]), _dev(s[:, 2]), _dev(c["q"][c["sorder"]]), _dev(t[:, 0]), _dev(t[:, 1]),
            _dev(t[:, 2])]


def _zeroed(n):
    import torch

    t = torch.full((n,), 7.0, dtype=torch.float64, device="cuda:0")
    _lib.check(_lib.lib().fmmh_kernel_zero(0, _ptr(t), n, None))
    return t


def test_p2p_kernel(case):
    c = case
    print("case kernel p2p", flush=True)
    ref = _run_operator(c, [("p2p", range(c["L"]))], "output")
    off, ent = _csr(c["plan"], "near", range(c["L"]))
    keep = [_dev(off), _dev(ent), _dev(c["cells"]), _dev(c["sstart"]), _dev(c["tstart"])] + _points(c)
    a = _lib.P2pInputs(0, len(c["cells"]), *[t.data_ptr() for t in keep])
    nt = len(c["pts"])
    phi_hat, grad_hat = _zeroed(nt), _zeroed(3 * nt)
    _lib.check(_lib.lib().fmmh_kernel_p2p(0, C.byref(a), 1, _ptr(phi_hat), _ptr(grad_hat), None))
    phi, grad = _gather(c, phi_hat, grad_hat)
    assert np.array_equal(phi, ref.potential) and np.array_equal(grad, ref.gradient)
    # the near rows against the oracle's direct sum, for a few target leaves
    pts, q, sp = c["pts"], c["q"], c["sstart"]
    for lf in range(0, len(c["cells"]), max(1, len(c["cells"]) // 7)):
        t = c["torder"][c["tstart"][lf]:c["tstart"][lf + 1]]
        if len(t) == 0:
            continue
        src = np.concatenate([c["sorder"][sp[s]:sp[s + 1]] for s in ent[off[lf]:off[lf + 1]]])
        dp, _ = O.direct_sum(pts[src], q[src], pts[t])
        assert np.allclose(phi[t], dp / (4 * np.pi), rtol=1e-12, atol=1e-12 * np.abs(dp).max())


def test_p2m_and_p2l_kernels(case):
    import torch

    c = case
    print("case kernel p2m p2l", flush=True)
    nc, nbt = c["nc"], int(c["box_base"][-1])
    pt = _points(c)
    # P2M: one row per leaf with sources, its own box
    leaves = np.flatnonzero(np.diff(c["sstart"]) > 0)
    rows = [_dev(c["leaf_box"][leaves].astype(np.uint32)), _dev(c["cells"][leaves]),
            _dev(np.arange(len(leaves) + 1, dtype=np.uint32)), _dev(leaves.astype(np.uint32)), _dev(c["cells"]),
            _dev(c["sstart"])]
    a = _lib.SourceInputs(len(leaves), *[t.data_ptr() for t in rows + pt[:4]])
    coeffs = _zeroed(nbt * nc)
    _lib.check(_lib.lib().fmmh_kernel_p2m(0, c["p"], C.byref(a), _ptr(coeffs), None))
    torch.cuda.synchronize()
    ref = _run_operator(c, [("p2m", range(c["L"]))], "multipoles")
    assert np.array_equal(coeffs.cpu().numpy().reshape(nbt, nc), ref)
    # P2L: the X rows of every level's boxes
    off, ent = _csr(c["plan"], "x", range(c["L"]))
    ids = np.arange(nbt, dtype=np.uint32)
    rows = [_dev(ids), _dev(c["box_cells"]), _dev(off), _dev(ent), _dev(c["cells"]), _dev(c["sstart"])]
    a = _lib.SourceInputs(nbt, *[t.data_ptr() for t in rows + pt[:4]])
    coeffs = _zeroed(nbt * nc)
    _lib.check(_lib.lib().fmmh_kernel_p2l(0, c["p"], C.byref(a), _ptr(coeffs), None))
    torch.cuda.synchronize()
    ref = _run_operator(c, [("p2l", range(c["L"]))], "locals")
    assert len(ent) > 1 and np.array_equal(coeffs.cpu().numpy().reshape(nbt, nc), ref)


def test_l2p_and_m2p_kernels(case):
    c = case
    print("case kernel l2p m2p", flush=True)
    nt, L = len(c["pts"]), c["L"]
    down = [("p2m", range(L)), ("m2m", range(L - 1, -1, -1)), ("m2l", range(L)), ("p2l", range(L)),
            ("l2l", range(L))]
    pt = _points(c)
    # L2P: the locals after the downward pass, at every leaf's targets
    loc = _run_operator(c, down, "locals")
    ref = _run_operator(c, down + [("l2p", range(L))], "output")
    leaves = np.flatnonzero(np.diff(c["tstart"]) > 0).astype(np.uint32)
    rows = [_dev(leaves), _dev(np.arange(len(leaves) + 1, dtype=np.uint32)),
            _dev(c["leaf_box"][leaves].astype(np.uint32)), _dev(c["box_cells"]), _dev(c["cells"]),
            _dev(c["tstart"])]
    a = _lib.TargetInputs(len(leaves), *[t.data_ptr() for t in rows + pt[4:]])
    phi_hat, grad_hat = _zeroed(nt), _zeroed(3 * nt)
    _lib.check(_lib.lib().fmmh_kernel_l2p(0, c["p"], 1, C.byref(a), _ptr(_dev(loc.ravel())), _ptr(phi_hat),
                                          _ptr(grad_hat), None))
    phi, grad = _gather(c, phi_hat, grad_hat)
    assert np.array_equal(phi, ref.potential) and np.array_equal(grad, ref.gradient)
    # M2P: the multipoles of the upward pass at the W-list boxes of every leaf
    up = [("p2m", range(L)), ("m2m", range(L - 1, -1, -1))]
    mult = _run_operator(c, up, "multipoles")
    ref = _run_operator(c, up + [("m2p", range(L))], "output")
    bb = c["box_base"]
    off, ent = _csr(c["plan"], "w", range(L), lambda l, e: (np.asarray(e, np.int64) + (bb[l + 1] if l + 1 < L else 0)))
    rows = [_dev(np.arange(len(c["cells"]), dtype=np.uint32)), _dev(off), _dev(ent), _dev(c["box_cells"]),
            _dev(c["cells"]), _dev(c["tstart"])]
    a = _lib.TargetInputs(len(c["cells"]), *[t.data_ptr() for t in rows + pt[4:]])
    phi_hat, grad_hat = _zeroed(nt), _zeroed(3 * nt)
    _lib.check(_lib.lib().fmmh_kernel_m2p(0, c["p"], 1, C.byref(a), _ptr(_dev(mult.ravel())), _ptr(phi_hat),
                                          _ptr(grad_hat), None))
    phi, grad = _gather(c, phi_hat, grad_hat)
    assert len(ent) > 1 and np.array_equal(phi, ref.potential) and np.array_equal(grad, ref.gradient)


def test_kernel_entry_points_refuse_null_inputs():
    a = _lib.P2pInputs()
    assert _lib.lib().fmmh_kernel_p2p(0, C.byref(a), 1, None, None, None) != 0
    assert _lib.lib().fmmh_kernel_zero(0, None, 5, None) != 0
