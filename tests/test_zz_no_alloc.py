"""fmmh_direct_sum_noalloc: the compensated direct sum with its partial sums in a fixed
per-GPU array (no allocation), against fmmh_direct_sum on the same device buffers.

Two calls in a row, the second smaller (the array keeps the first call's contents), each
compared bit for bit with the allocating entry point (the same source chunks), and the
first also with the oracle's compensated sum.  A call above the array's target bound is
refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib
from oracle import oracle as O


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _case(torch, ns, nt, seed):
    rng = O.Rng(seed)
    src = rng.cube(ns)
    q = rng.normals(ns)
    tgt = np.vstack([src[: nt // 2], rng.cube(nt - nt // 2)])  # half the targets coincide with sources
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    return src, q, tgt, t(src), t(q), t(tgt)


def _noalloc(torch, s, q, t, ns, nt):
    dev = torch.device("cuda", 0)
    phi = torch.full((nt,), np.nan, dtype=torch.float64, device=dev)
    grad = torch.full((nt, 3), np.nan, dtype=torch.float64, device=dev)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(_lib.lib().fmmh_direct_sum_noalloc(0, _ptr(s), _ptr(q), ns, _ptr(t), nt, _ptr(phi), _ptr(grad),
                                                  stream))
    torch.cuda.synchronize()
    return phi.cpu().numpy(), grad.cpu().numpy()


def _alloc(torch, s, q, t, ns, nt):
    dev = torch.device("cuda", 0)
    phi = torch.zeros(nt, dtype=torch.float64, device=dev)
    grad = torch.zeros(nt, 3, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().fmmh_direct_sum(0, _ptr(s), _ptr(q), ns, _ptr(t), nt, _ptr(phi), _ptr(grad), 1))
    return phi.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_direct_sum_matches_allocating_entry_point(torch_gpu):
    torch = torch_gpu
    # first call: 40,000 sources, 5,000 targets (20 target blocks, 40 source chunks)
    src, q, tgt, s_d, q_d, t_d = _case(torch, 40000, 5000, 91)
    phi, grad = _noalloc(torch, s_d, q_d, t_d, 40000, 5000)
    rphi, rgrad = _alloc(torch, s_d, q_d, t_d, 40000, 5000)
    assert np.array_equal(phi, rphi) and np.array_equal(grad, rgrad)
    ophi, ograd = O.direct_sum(src, q, tgt)
    assert np.abs(phi - ophi).max() <= 1e-12 * np.abs(ophi).max()
    assert np.abs(grad - ograd).max() <= 1e-12 * np.abs(ograd).max()
    # second call, smaller, on the same array: 7,000 sources, 300 targets (2 target
    # blocks, 7 source chunks)
    src, q, tgt, s_d, q_d, t_d = _case(torch, 7000, 300, 92)
    phi, grad = _noalloc(torch, s_d, q_d, t_d, 7000, 300)
    rphi, rgrad = _alloc(torch, s_d, q_d, t_d, 7000, 300)
    assert np.array_equal(phi, rphi) and np.array_equal(grad, rgrad)
    ophi, ograd = O.direct_sum(src, q, tgt)
    assert np.abs(phi - ophi).max() <= 1e-12 * np.abs(ophi).max()


@pytest.mark.gpu
def test_noalloc_direct_sum_refuses_too_many_targets(torch_gpu):
    torch = torch_gpu
    dev = torch.device("cuda", 0)
    n = 262145
    buf = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    code = _lib.lib().fmmh_direct_sum_noalloc(0, _ptr(buf), _ptr(buf), 1, _ptr(buf), n, _ptr(buf), None, None)
    assert code == 5  # FMMH_TOO_LARGE, nothing launched
