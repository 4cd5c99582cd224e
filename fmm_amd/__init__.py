"""fmm_amd -- the FMM evaluation hot path of tbetcke/fmm on one MI355X.

Host mirror of the reference's evaluate/plan API (fmm-exec/src/fmm.rs):

    fmm = FmmBuilder(p=8).gradients(True).max_points_per_leaf(128).build(sources, targets)
    out = fmm.evaluate(charges)        # Output(potential, gradient)

The octree and interaction lists are built on the CPU by the library's C++ host code
(the role of octree + fmm-plan); every operator of the evaluation runs as hand-written
HIP kernels for gfx950 behind the C ABI of include/fmm_amd.h.  There is no CPU
evaluation path: without the built library or a GPU the calls raise.
"""
from __future__ import annotations

import copy
import ctypes as C
import os
from dataclasses import dataclass
from typing import NamedTuple, Optional

import numpy as np

from . import _lib
from ._lib import FmmError, build  # noqa: F401

PLAIN_F64 = 0
MIXED = 1
PRECISIONS = {"f64": PLAIN_F64, "plain": PLAIN_F64, "mixed": MIXED}
M2L_METHODS = {"dense": 0, "rotation": 1, "plane_wave": 2}
KINDS = ("p2m", "m2m", "l2l", "m2l", "p2l", "l2p", "m2p", "p2p", "other", "m2l_contract")
LISTS = {"near": 0, "v": 1, "w": 2, "x": 3}


def _device_buffer(t, count: int, name: str, device: int, complex_=False):
    """The data pointer of a caller's device tensor after checking it against what the
    operator reads or writes: float64 (or complex128 for complex values), contiguous,
    `count` elements, on the operator's GPU.  Raises FmmError(InvalidArgument) else."""
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise FmmError(9, f"{name}: expected a CUDA tensor")
    want = torch.complex128 if complex_ else torch.float64
    if t.dtype != want:
        raise FmmError(9, f"{name}: dtype {t.dtype}, expected {want}")
    if not t.is_contiguous():
        raise FmmError(9, f"{name}: not contiguous")
    if t.numel() != count:
        raise FmmError(9, f"{name}: {t.numel()} elements, expected {count}")
    if t.device.index != device:
        raise FmmError(6, f"{name}: on cuda:{t.device.index}, the operator runs on cuda:{device}")
    return C.c_void_p(t.data_ptr())


def _torch_stream(device: int):
    """The caller's current torch stream on `device` (a hipStream_t), for fmmh_stream_wait."""
    import torch

    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def device_count() -> int:
    n = C.c_int32(0)
    _lib.check(_lib.lib().fmmh_device_count(C.byref(n)))
    return int(n.value)


def set_table_cache(directory) -> None:
    """Keep the Laplace tables (M2M, L2L, dense M2L) of every degree built from now on in
    `directory`, and load them from there when valid (the reference's TableCache file
    format, fmm-tables/src/cache.rs); None: build every time."""
    d = None if directory is None else os.fsencode(os.fspath(directory))
    if d is not None:
        os.makedirs(d, exist_ok=True)
    _lib.check(_lib.lib().fmmh_set_table_cache(d))


def table_cache_stats():
    """(hits, misses) of the table cache since it was set."""
    h, m = C.c_int64(0), C.c_int64(0)
    _lib.check(_lib.lib().fmmh_table_cache_stats(C.byref(h), C.byref(m)))
    return h.value, m.value


def tables(p: int):
    """(M2M[8], L2L[8], M2L[316]) matrices A, output = A @ input (CONVENTIONS §3.12)."""
    n = (p + 1) ** 2
    a, b, c = np.zeros(8 * n * n), np.zeros(8 * n * n), np.zeros(316 * n * n)
    _lib.check(_lib.lib().fmmh_tables(p, _lib.ptr(a), _lib.ptr(b), _lib.ptr(c)))
    f = lambda x, k: x.reshape(k, n, n).transpose(0, 2, 1).copy()
    return f(a, 8), f(b, 8), f(c, 316)


class Plan:
    """Tree, box index, lists and leaf-ordered points (one rank), built on the CPU, or with
    device set its per-point stages (bounding box, Morton keys and their sort, the points'
    leaves, order and coordinates) on that GPU: the same plan bit for bit."""

    def __init__(self, sources, targets=None, leaf_points=64, max_level=16, near_rule=None, device=None):
        src = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
        tgt = src if targets is None else np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
        self.sources, self.targets = src, tgt
        o = _lib.PlanOptions(int(leaf_points), int(max_level), -1.0, -1.0)
        if near_rule is not None:
            o.near_w, o.near_x = float(near_rule[0]), float(near_rule[1])
        h = C.c_void_p()
        if device is None:
            _lib.check(_lib.lib().fmmh_plan_build(_lib.ptr(src), len(src), _lib.ptr(tgt), len(tgt), C.byref(o),
                                                  C.byref(h)))
        else:
            _lib.check(_lib.lib().fmmh_plan_build_device(int(device), _lib.ptr(src), len(src), _lib.ptr(tgt), len(tgt),
                                                         C.byref(o), C.byref(h)))
        self._h = h
        self._free = _lib.lib().fmmh_plan_free

    @classmethod
    def forest(cls, clouds, leaf_points=64, max_level=16, near_rule=None):
        """Independent clouds (a list of (n_c, 3) arrays; sources = targets in each) as one
        plan: each cloud's own tree, every cloud evaluated by the same launches."""
        arrays = [np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3) for c in clouds]
        pts = np.ascontiguousarray(np.vstack(arrays))
        counts = np.array([len(a) for a in arrays], dtype=np.int64)
        o = _lib.PlanOptions(int(leaf_points), int(max_level), -1.0, -1.0)
        if near_rule is not None:
            o.near_w, o.near_x = float(near_rule[0]), float(near_rule[1])
        self = cls.__new__(cls)
        self.sources = self.targets = pts
        self.counts = counts
        h = C.c_void_p()
        _lib.check(_lib.lib().fmmh_forest_build(_lib.ptr(pts), _lib.ptr(counts), len(counts), C.byref(o),
                                                C.byref(h)))
        self._h = h
        self._free = _lib.lib().fmmh_plan_free
        return self

    def __del__(self):
        # the library's free function held by the object: at interpreter exit the module's
        # globals may already be gone
        h = getattr(self, "_h", None)
        free = getattr(self, "_free", None)
        if h and free is not None:
            free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def nlevels(self) -> int:
        return _lib.lib().fmmh_plan_nlevels(self._h)

    @property
    def nleaves(self) -> int:
        return _lib.lib().fmmh_plan_nleaves(self._h)

    def domain(self):
        box, side = np.zeros(6), C.c_double()
        _lib.check(_lib.lib().fmmh_plan_domain(self._h, _lib.ptr(box), C.byref(side)))
        return box, side.value

    def box_keys(self, level: int) -> np.ndarray:
        out = np.zeros(_lib.lib().fmmh_plan_nboxes(self._h, level), np.uint64)
        if len(out):
            _lib.check(_lib.lib().fmmh_plan_box_keys(self._h, level, _lib.ptr(out)))
        return out

    def leaf_keys(self) -> np.ndarray:
        out = np.zeros(self.nleaves, np.uint64)
        _lib.check(_lib.lib().fmmh_plan_leaf_keys(self._h, _lib.ptr(out)))
        return out

    def list(self, kind: str, level: int):
        k = LISTS[kind]
        L = _lib.lib()
        total = L.fmmh_plan_list(self._h, k, level, None, None, None)
        if total < 0:
            raise FmmError(9, L.fmmh_last_error().decode())
        if kind in ("near", "w"):
            rows = int(np.sum((self.leaf_keys() & np.uint64(0x7FFF)) == level))
        else:
            rows = len(self.box_keys(level))
        off = np.zeros(rows + 1, np.uint32)
        ent = np.zeros(max(total, 1), np.uint32)
        aux = np.zeros(max(total, 1), np.uint16)
        L.fmmh_plan_list(self._h, k, level, _lib.ptr(off), _lib.ptr(ent), _lib.ptr(aux))
        if kind == "v":
            return off, ent[:total], aux[:total]
        return off, ent[:total]

    def point_leaves(self, targets=False) -> np.ndarray:
        n = len(self.targets) if targets else len(self.sources)
        out = np.zeros(n, np.uint32)
        _lib.check(_lib.lib().fmmh_plan_point_leaves(self._h, int(targets), _lib.ptr(out)))
        return out

    def leaf_coordinates(self, targets=False):
        n = len(self.targets) if targets else len(self.sources)
        u, order = np.zeros((n, 3)), np.zeros(n, np.uint32)
        _lib.check(_lib.lib().fmmh_plan_leaf_coordinates(self._h, int(targets), _lib.ptr(u), _lib.ptr(order)))
        return u, order

    def leaf_cells(self) -> np.ndarray:
        """(level, ix, iy, iz) of every leaf, uint32 [nleaves, 4]."""
        out = np.zeros((self.nleaves, 4), np.uint32)
        _lib.check(_lib.lib().fmmh_plan_leaf_cells(self._h, _lib.ptr(out)))
        return out

    def box_cells(self, level: int) -> np.ndarray:
        """(level, ix, iy, iz) of every box of a level, uint32 [nboxes, 4]."""
        out = np.zeros((len(self.box_keys(level)), 4), np.uint32)
        _lib.check(_lib.lib().fmmh_plan_box_cells(self._h, int(level), _lib.ptr(out)))
        return out

    def point_starts(self, targets=False) -> np.ndarray:
        """CSR of each leaf's points in leaf order (nleaves + 1 offsets)."""
        out = np.zeros(self.nleaves + 1, np.uint32)
        _lib.check(_lib.lib().fmmh_plan_point_starts(self._h, int(targets), _lib.ptr(out)))
        return out


@dataclass
class Output:
    """fmm-exec's Output: potential[nt] and, with gradients, gradient[nt, 3]."""

    potential: np.ndarray
    gradient: Optional[np.ndarray]


class Operator:
    """The device operator: the level calls of FmmOperator (fmm-plan/src/operator.rs)
    on the GPU, plus the evaluator's whole order in one call."""

    def __init__(self, plan: Plan, p: int, precision="f64", gradients=True, device=0, m2l="dense", tuning=None):
        self.plan = plan
        self.p = p
        self.gradients = gradients
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.m2l = M2L_METHODS[m2l] if isinstance(m2l, str) else int(m2l)
        self.device = int(device)
        o = _lib.OperatorOptions(int(p), self.precision, int(bool(gradients)), int(device), self.m2l)
        h = C.c_void_p()
        t = C.byref(tuning) if tuning is not None else None
        _lib.check(_lib.lib().fmmh_operator_create_tuned(plan.handle, C.byref(o), t, C.byref(h)))
        self._h = h
        self._free = _lib.lib().fmmh_operator_free
        self.tuning = tuning
        self.ntargets = len(plan.targets)
        self.nsources = len(plan.sources)

    def __del__(self):
        # the library's free function held by the object: at interpreter exit the module's
        # globals may already be gone
        h = getattr(self, "_h", None)
        free = getattr(self, "_free", None)
        if h and free is not None:
            free(h)
            self._h = None

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(_lib.lib().fmmh_operator_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_}

    def load_charges(self, charges):
        """Charges in caller order: a numpy array, or a float64 CUDA tensor on the
        operator's GPU (the operator's stream is ordered after torch's current stream)."""
        if hasattr(charges, "data_ptr") and getattr(charges, "is_cuda", False):
            ptr = _device_buffer(charges, self.nsources, "charges", self.device)
            _lib.check(_lib.lib().fmmh_stream_wait(self._h, _torch_stream(self.device)))
            _lib.check(_lib.lib().fmmh_load_charges(self._h, ptr, self.nsources, 1))
            return
        q = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
        if len(q) != self.nsources:
            raise FmmError(9, f"expected {self.nsources} charges, got {len(q)}")
        _lib.check(_lib.lib().fmmh_load_charges(self._h, _lib.ptr(q), len(q), 0))

    def reset(self):
        _lib.check(_lib.lib().fmmh_reset(self._h))

    def call(self, kind: str, level: int):
        _lib.check(getattr(_lib.lib(), "fmmh_" + kind)(self._h, int(level)))

    def evaluate(self):
        _lib.check(_lib.lib().fmmh_evaluate(self._h))

    def evaluate_graph(self):
        """reset() + evaluate() as one replayed hipGraph launch."""
        _lib.check(_lib.lib().fmmh_evaluate_graph(self._h))

    def read_output(self) -> Output:
        phi = np.zeros(self.ntargets)
        grad = np.zeros((self.ntargets, 3)) if self.gradients else None
        _lib.check(_lib.lib().fmmh_read_output(self._h, _lib.ptr(phi), _lib.ptr(grad), self.ntargets, 0))
        return Output(phi, grad)

    def read_output_device(self, phi, grad=None):
        """Output pass into device buffers (float64 CUDA tensors on the operator's GPU:
        phi of ntargets, grad of 3 ntargets); complete on return."""
        pp = _device_buffer(phi, self.ntargets, "phi", self.device)
        g = None
        if grad is not None:
            if not self.gradients:
                raise FmmError(7, "read_output_device: the operator computes no gradients")
            g = _device_buffer(grad, 3 * self.ntargets, "grad", self.device)
        _lib.check(_lib.lib().fmmh_stream_wait(self._h, _torch_stream(self.device)))
        _lib.check(_lib.lib().fmmh_read_output(self._h, pp, g, self.ntargets, 1))

    def sync(self):
        _lib.check(_lib.lib().fmmh_sync(self._h))

    def vjp(self, charges, cotangent, field=None):
        """dL/dq and (with the forward gradient `field`) dL/dpoints for a plan whose
        targets are its sources; numpy arrays in caller order."""
        q = np.ascontiguousarray(charges, dtype=np.float64)
        g = np.ascontiguousarray(cotangent, dtype=np.float64)
        gq = np.zeros(len(q))
        f = gp = None
        if field is not None:
            f = np.ascontiguousarray(field, dtype=np.float64).reshape(-1, 3)
            gp = np.zeros((len(q), 3))
        if len(q) != self.nsources or len(g) != self.nsources:
            raise FmmError(9, f"vjp: expected {self.nsources} charges and cotangents")
        if f is not None and len(f) != self.nsources:
            raise FmmError(9, f"vjp: expected a field of {self.nsources} x 3")
        _lib.check(_lib.lib().fmmh_vjp(self._h, _lib.ptr(q), _lib.ptr(g), _lib.ptr(f), len(q), 0, _lib.ptr(gq),
                                       _lib.ptr(gp)))
        return gq, gp

    def vjp_device(self, charges, cotangent, field, grad_charges, grad_points=None):
        """The same on device buffers (float64 CUDA tensors on the operator's GPU, n and
        3 n elements); the operator's stream waits for torch's current stream first, and
        the outputs are complete on return."""
        n, d = self.nsources, self.device
        args = [_device_buffer(charges, n, "charges", d), _device_buffer(cotangent, n, "cotangent", d),
                None if field is None else _device_buffer(field, 3 * n, "field", d)]
        outs = [_device_buffer(grad_charges, n, "grad_charges", d),
                None if grad_points is None else _device_buffer(grad_points, 3 * n, "grad_points", d)]
        _lib.check(_lib.lib().fmmh_stream_wait(self._h, _torch_stream(d)))
        _lib.check(_lib.lib().fmmh_vjp(self._h, *args, n, 1, *outs))

    def expansions(self, locals_: bool, level: int) -> np.ndarray:
        n = (self.p + 1) ** 2
        out = np.zeros((len(self.plan.box_keys(level)), n))
        _lib.check(_lib.lib().fmmh_read_expansions(self._h, int(locals_), int(level), _lib.ptr(out)))
        return out

    def kind_timing(self, enable=True):
        _lib.check(_lib.lib().fmmh_kind_timing(self._h, int(enable)))

    def kind_times(self):
        ms = np.zeros(len(KINDS))
        n = np.zeros(len(KINDS), np.int64)
        _lib.check(_lib.lib().fmmh_kind_times(self._h, _lib.ptr(ms), _lib.ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KINDS)}

    def inject_error(self):
        """Test hook (DeviceOperator::inject_error, fmm-exec/src/device.rs:4621): the next
        call that touches the device fails with an injected device error, which the
        operator keeps and returns from every such call after it."""
        _lib.check(_lib.lib().fmmh_inject_error(self._h))


def direct_sum(sources, charges, targets, gradients=True, device=0):
    """Compensated f64 direct sum on the GPU (unscaled: sum q / r, -sum q d / r^3), as
    nd_fmm_ref::p2p::direct_sum (fmm-ref/src/p2p.rs:147-205)."""
    src = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(charges, dtype=np.float64)
    tgt = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    phi = np.zeros(len(tgt))
    grad = np.zeros((len(tgt), 3)) if gradients else None
    _lib.check(_lib.lib().fmmh_direct_sum(int(device), _lib.ptr(src), _lib.ptr(q), len(src), _lib.ptr(tgt), len(tgt),
                                          _lib.ptr(phi), _lib.ptr(grad), 0))
    return phi, grad


class Fmm:
    """An FMM built for fixed sources and targets (fmm-exec's Fmm)."""

    def __init__(self, plan: Plan, operator: Operator):
        self.plan = plan
        self.operator = operator

    def inject_device_error(self):
        """Test hook (Fmm::inject_device_error, fmm-exec/src/fmm.rs:4715): the next
        evaluation fails with FmmError(FMMH_DEVICE), and so does every later one."""
        self.operator.inject_error()

    def evaluate(self, charges) -> Output:
        """Potential and gradient at the targets (1/(4 pi) included), caller order."""
        op = self.operator
        op.load_charges(charges)
        op.reset()
        op.evaluate()
        return op.read_output()

    def evaluate_levels(self, charges) -> Output:
        """The same evaluation issued as the reference's level calls, one by one
        (fmm-plan/src/operator.rs:408-425)."""
        op, L = self.operator, self.plan.nlevels - 1
        op.load_charges(charges)
        op.reset()
        for level in range(L, -1, -1):
            op.call("p2m", level)
            if level >= 1:
                op.call("m2m", level - 1)
        for level in range(1, L + 1):
            for k in ("l2l", "m2l", "p2l"):
                op.call(k, level)
        for level in range(0, L + 1):
            for k in ("l2p", "m2p", "p2p"):
                op.call(k, level)
        return op.read_output()

    def vjp(self, charges, cotangent, field=None):
        """Adjoint of evaluate's potentials (sources = targets): (dL/dq, dL/dpoints) for the
        cotangent dL/dphi; `field` is evaluate's gradient (needed for dL/dpoints)."""
        return self.operator.vjp(charges, cotangent, field)


class FmmBuilder:
    """Settings of an FMM (fmm-exec/src/fmm.rs:1261-1760), built with build()."""

    def __init__(self, p: int = 6):
        self._p = p
        self._gradients = True
        self._max_level = 16
        self._leaf_points = 64
        self._precision = "f64"
        self._device = 0
        self._near = None
        self._m2l = "dense"
        self._tuning_dir = None
        self._tuning_budget = 10.0
        self._leaf_candidates = None
        self.tuning_info = None

    def gradients(self, on: bool):
        self._gradients = bool(on)
        return self

    def tuning_cache(self, directory):
        """Autotune the operator's launch shapes at build and keep the decisions in
        <directory>/fmm_amd_tuning.json (FmmBuilder::tuning_cache, fmm-exec/src/tune.rs);
        None: the static rule, nothing timed."""
        self._tuning_dir = None if directory is None else os.fspath(directory)
        return self

    def tuning_budget(self, seconds: float):
        """Seconds the autotune may spend per build (10 by default, as the reference's)."""
        self._tuning_budget = float(seconds)
        return self

    def tune_leaf_size(self, candidates):
        """Choose max_points_per_leaf among `candidates` by timing whole evaluations of
        the points given to build (cached per N bucket with tuning_cache)."""
        self._leaf_candidates = tuple(int(c) for c in candidates) if candidates else None
        return self

    def _with_leaf(self, leaf: int):
        b = copy.copy(self)
        b._leaf_points = int(leaf)
        return b

    def max_level(self, level: int):
        self._max_level = int(level)
        return self

    def max_points_per_leaf(self, n: int):
        self._leaf_points = int(n)
        return self

    def precision(self, precision: str):
        if precision not in PRECISIONS:
            raise FmmError(3, f"unknown precision {precision!r}")
        self._precision = precision
        return self

    def device(self, ordinal: int):
        self._device = int(ordinal)
        return self

    def near_rule(self, rule):
        """The near rule (fmm-exec/src/fmm.rs:1018-1050): None or "off"; (N_W, N_X) named
        crossovers (NearRuleChoice::Crossovers); or "device", the crossovers of this
        GPU path's cost ratio, N_W = N_X = (p+1)^2 / 2 points (an M2P or P2L costs about
        (p+1)^2 / 2 pair interactions per point on the device)."""
        if rule is None or rule == "off":
            self._near = None
        elif rule == "device":
            c = 0.5 * (self._p + 1) ** 2
            self._near = (c, c)
        else:
            self._near = (float(rule[0]), float(rule[1]))
        return self

    def m2l(self, method: str):
        """The M2L formulation (fmm-exec's M2lMethod): "dense" (tables on MFMA), "rotation"
        (rotate, coaxial shift, rotate back; f32 in the mixed precision) or "plane_wave"
        (waves in six directions by merge and shift, f64 in both precisions, on FMM3D's
        quadrature of the degree: its error about a tenth of the truncation's)."""
        if method not in M2L_METHODS:
            raise ValueError(f"unknown M2L method {method!r}")
        self._m2l = method
        return self

    def build_plan(self, sources, targets=None) -> Plan:
        return Plan(sources, targets, self._leaf_points, self._max_level, self._near)

    def _operator(self, plan) -> Operator:
        tuning = None
        if self._tuning_dir is not None:
            from . import tune

            tuning, self.tuning_info = tune.resolve(plan, self._p, self._precision, self._gradients, self._device,
                                                    self._m2l, self._tuning_dir, self._tuning_budget)
        return Operator(plan, self._p, self._precision, self._gradients, self._device, self._m2l, tuning)

    def build(self, sources, targets=None) -> Fmm:
        b = self
        if self._leaf_candidates and targets is None:
            from . import tune

            leaf, times = tune.tune_leaf_size(self, sources, self._leaf_candidates, self._tuning_dir)
            b = self._with_leaf(leaf)
            self.leaf_info = {"leaf_points": leaf, "ms": times}
        plan = b.build_plan(sources, targets)
        op = b._operator(plan)
        self.tuning_info = b.tuning_info
        return Fmm(plan, op)

    def build_batched(self, clouds) -> Fmm:
        """One Fmm over many independent clouds (sources = targets in each); charges and
        outputs are the clouds' points in order."""
        plan = Plan.forest(clouds, self._leaf_points, self._max_level, self._near)
        return Fmm(plan, self._operator(plan))


class PlaneWaveQuadrature(NamedTuple):
    """FMM3D's plane-wave quadrature as the reference uses it (quadrature.rs Quadrature)."""
    lambdas: np.ndarray
    weights: np.ndarray
    alpha_nodes: np.ndarray  # FMM3D's N_k (odd counts included)
    fourier: np.ndarray      # modes kept at degree p: min(F_k, p + 1)
    half_terms: int          # H = sum_k N_k' / 2, N_k' = N_k rounded up to even
    max_error: float         # FMM3D's stated maximum error


def plane_wave_quadrature(p: int, nodes: int = 0) -> PlaneWaveQuadrature:
    """The plane-wave quadrature of degree p: FMM3D's table with ``nodes`` lambda nodes, or
    the reference's count for degree p when nodes is 0."""
    n, h, err = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
    lam, w = np.zeros(40), np.zeros(40)
    na, fo = np.zeros(40, np.int32), np.zeros(40, np.int32)
    _lib.check(_lib.lib().fmmh_plane_wave_quadrature(int(p), int(nodes), C.byref(n), C.byref(h), _lib.ptr(lam),
                                                     _lib.ptr(w), _lib.ptr(na), _lib.ptr(fo), C.byref(err)))
    k = n.value
    return PlaneWaveQuadrature(lam[:k].copy(), w[:k].copy(), na[:k].copy(), fo[:k].copy(), h.value, err.value)


def plane_wave_m2l(p: int, offset: int, multipole, direction: int = -1, nodes: int = 0) -> np.ndarray:
    """One V pair's plane-wave M2L on the host (scaled storage): the local it adds, in the
    given direction (-1: the offset's own)."""
    x = np.ascontiguousarray(multipole, dtype=np.float64)
    out = np.zeros((p + 1) ** 2)
    _lib.check(_lib.lib().fmmh_plane_wave_m2l(int(p), int(nodes), int(offset), int(direction), _lib.ptr(x),
                                              _lib.ptr(out)))
    return out


def plane_wave_merge(p: int, slot_multipoles, present, nodes: int = 0) -> np.ndarray:
    """One parent's merge and shift on the host: 216 slot multipoles (rows), which are held,
    and the 8 children's locals (rows) it adds."""
    x = np.ascontiguousarray(slot_multipoles, dtype=np.float64).reshape(216, (p + 1) ** 2)
    pr = np.ascontiguousarray(present, dtype=np.uint8).reshape(216)
    out = np.zeros((8, (p + 1) ** 2))
    _lib.check(_lib.lib().fmmh_plane_wave_merge(int(p), int(nodes), _lib.ptr(x), _lib.ptr(pr), _lib.ptr(out)))
    return out


def plane_wave_slot_direction(slot: int, child: int) -> int:
    """The direction (0..5) of the merge-and-shift pair slot -> child, -1 if adjacent."""
    return int(_lib.lib().fmmh_plane_wave_slot_direction(int(slot), int(child)))


# ---- Helmholtz --------------------------------------------------------------------------

def helmholtz_degree(k, radius: float, tolerance: float) -> int:
    """The degree rule of fmm-exec/src/helmholtz/degree.rs:8-15 for boxes of half-width
    `radius`: max(ceil(0.30 L^2 + 0.25 L + 2.60), ceil(x + 0.80 (ln 1/eps)^(2/3) x^(1/3))),
    x = Re(k) sqrt(3) radius, L = log10(1/eps)."""
    p = _lib.lib().fmmh_helmholtz_degree(float(np.real(k)), float(radius), float(tolerance))
    if p < 0:
        raise FmmError(9, "helmholtz_degree: 0 < tolerance < 1, radius > 0, Re k >= 0")
    return int(p)


def helmholtz_table(p: int, kh, kind: str, index: int) -> np.ndarray:
    """One translation table of a level whose wavenumber in its own units is kh: "m2m" from
    the child of octant `index`, "l2l" into it, "m2l" of offset index `index`, "m2l_rcr" the
    same M2L composed from the per-pair rotation path's pieces (rotation blocks, scaled
    coaxial blocks, rotation back); a complex matrix A with stored output = A @ stored input."""
    n = (p + 1) ** 2
    out = np.zeros(2 * n * n)
    k = {"m2m": 0, "l2l": 1, "m2l": 2, "m2l_rcr": 3}[kind]
    _lib.check(_lib.lib().fmmh_helmholtz_table(int(p), float(np.real(kh)), float(np.imag(kh)), k, int(index),
                                               _lib.ptr(out)))
    return (out[: n * n] + 1j * out[n * n:]).reshape(n, n).T.copy()


def helmholtz_direct_sum(k, sources, charges, targets, gradients=True, device=0):
    """Compensated f64 direct sum on the GPU of q e^{ikr}/r and its gradient (no 1/(4 pi)),
    nd_fmm_ref::helmholtz::p2p::direct_sum (fmm-ref/src/helmholtz/p2p.rs:155); complex."""
    src = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.complex128).view(np.float64))
    tgt = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    phi = np.zeros(2 * len(tgt))
    grad = np.zeros(6 * len(tgt)) if gradients else None
    _lib.check(_lib.lib().fmmh_helmholtz_direct_sum(int(device), float(np.real(k)), float(np.imag(k)),
                                                    _lib.ptr(src), _lib.ptr(q), len(src), _lib.ptr(tgt), len(tgt),
                                                    _lib.ptr(phi), _lib.ptr(grad), 0))
    return phi.view(np.complex128), (grad.view(np.complex128).reshape(-1, 3) if gradients else None)


class HelmholtzPlaneWaveQuadrature(NamedTuple):
    """FMM3D's iprec-2 quadrature of kappa = k w as the reference maps it (quadrature.rs)."""
    domain: int              # 1-based, 0 outside the tables
    stated_error: float
    lambdas: np.ndarray      # complex
    t: np.ndarray            # complex, sqrt(lambda^2 + kappa^2)
    weights: np.ndarray      # complex
    alpha_counts: np.ndarray
    fourier_counts: np.ndarray


def helmholtz_plane_wave_quadrature(kappa) -> HelmholtzPlaneWaveQuadrature:
    """The exponential M2L's quadrature of kappa = k w (box widths)."""
    cap = 128
    dom, n, err = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
    lam, t, w = np.zeros(2 * cap), np.zeros(2 * cap), np.zeros(2 * cap)
    al, fo = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    _lib.check(_lib.lib().fmmh_helmholtz_plane_wave_quadrature(
        float(np.real(kappa)), float(np.imag(kappa)), cap, C.byref(dom), C.byref(n), C.byref(err), _lib.ptr(lam),
        _lib.ptr(t), _lib.ptr(w), _lib.ptr(al), _lib.ptr(fo)))
    k = n.value
    c = lambda v: v[: 2 * k].view(np.complex128).copy()
    return HelmholtzPlaneWaveQuadrature(dom.value, err.value, c(lam), c(t), c(w), al[:k].copy(), fo[:k].copy())


def bessel_j_norms(z, top: int) -> np.ndarray:
    """|J_0(z)| .. |J_top(z)| of complex z (Miller's recurrence, normalised by e^{iz})."""
    out = np.zeros(top + 1)
    _lib.check(_lib.lib().fmmh_bessel_j_norms(float(np.real(z)), float(np.imag(z)), int(top), _lib.ptr(out)))
    return out


def helmholtz_plane_wave_m2l(p: int, kh, offset: int, multipole, direction: int = -1) -> np.ndarray:
    """One V pair's exponential M2L on the host for a level with kh = k r_l: the stored
    local (complex, (p+1)^2) it adds to a stored multipole (complex)."""
    n = (p + 1) ** 2
    m = np.asarray(multipole, dtype=np.complex128).reshape(n)
    x = np.ascontiguousarray(np.concatenate([m.real, m.imag]))
    out = np.zeros(2 * n)
    _lib.check(_lib.lib().fmmh_helmholtz_plane_wave_m2l(int(p), float(np.real(kh)), float(np.imag(kh)), int(offset),
                                                        int(direction), _lib.ptr(x), _lib.ptr(out)))
    return out[:n] + 1j * out[n:]


def helmholtz_plane_wave_merge(p: int, kh, slot_multipoles, present) -> np.ndarray:
    """One parent's merge and shift on the host: 216 complex slot multipoles (rows), which
    are held; the 8 children's complex locals (rows) it adds."""
    n = (p + 1) ** 2
    m = np.asarray(slot_multipoles, dtype=np.complex128).reshape(216, n)
    x = np.ascontiguousarray(np.concatenate([m.real, m.imag], axis=1))
    pr = np.ascontiguousarray(present, dtype=np.uint8).reshape(216)
    out = np.zeros((8, 2 * n))
    _lib.check(_lib.lib().fmmh_helmholtz_plane_wave_merge(int(p), float(np.real(kh)), float(np.imag(kh)),
                                                          _lib.ptr(x), _lib.ptr(pr), _lib.ptr(out)))
    return out[:, :n] + 1j * out[:, n:]


class HelmholtzOperator:
    """The Helmholtz device operator over a plan (complex charges, kernel e^{ikr}/(4 pi r))."""

    def __init__(self, plan: Plan, p: int, k, precision="f64", gradients=True, device=0, m2l="dense",
                 dense_budget=0):
        """m2l (HelmholtzM2l, fmm-exec/src/helmholtz/operator.rs:448-479): "dense", one
        complex table per offset on the MFMA contraction on the levels whose tables fit
        dense_budget bytes (316 x 2 (p+1)^4 f64 per level, from the finest level up; 0 the
        reference's 1 GiB, negative no limit) and per-pair rotation on the others,
        "rotation", per pair rotation, coaxial shift, rotation back in f64 (RCR) on every
        level, with tables O(p^3) per distance, or "plane_wave", the exponential M2L by merge
        and shift in f64 on the levels the reference's rule gives a quadrature (degree <= 15,
        levels >= 2 whose kappa = k w lies in FMM3D's tables; HelmholtzM2l::PlaneWave) and
        per-pair rotation on the others."""
        if m2l not in ("dense", "rotation", "plane_wave"):
            raise FmmError(7, f"Helmholtz M2L: dense, rotation or plane_wave, not {m2l!r}")
        self.plan = plan
        self.p = p
        self.k = complex(k)
        self.m2l = m2l
        self.gradients = gradients
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.device = int(device)
        o = _lib.HelmholtzOptions(int(p), self.k.real, self.k.imag, self.precision, int(bool(gradients)),
                                  int(device), M2L_METHODS[m2l], int(dense_budget))
        h = C.c_void_p()
        _lib.check(_lib.lib().fmmh_helmholtz_create(plan.handle, C.byref(o), C.byref(h)))
        self._h = h
        self._free = _lib.lib().fmmh_helmholtz_free
        self.ntargets = len(plan.targets)
        self.nsources = len(plan.sources)

    def __del__(self):
        # the library's free function held by the object: at interpreter exit the module's
        # globals may already be gone
        h = getattr(self, "_h", None)
        free = getattr(self, "_free", None)
        if h and free is not None:
            free(h)
            self._h = None

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(_lib.lib().fmmh_helmholtz_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_}

    def load_charges(self, charges):
        if hasattr(charges, "data_ptr") and getattr(charges, "is_cuda", False):
            ptr = _device_buffer(charges, self.nsources, "charges", self.device, complex_=True)
            _lib.check(_lib.lib().fmmh_helmholtz_stream_wait(self._h, _torch_stream(self.device)))
            _lib.check(_lib.lib().fmmh_helmholtz_load_charges(self._h, ptr, self.nsources, 1))
            return
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.complex128).reshape(-1))
        if len(q) != self.nsources:
            raise FmmError(9, f"expected {self.nsources} charges, got {len(q)}")
        _lib.check(_lib.lib().fmmh_helmholtz_load_charges(self._h, _lib.ptr(q.view(np.float64)), len(q), 0))

    def reset(self):
        _lib.check(_lib.lib().fmmh_helmholtz_reset(self._h))

    def evaluate(self):
        _lib.check(_lib.lib().fmmh_helmholtz_evaluate(self._h))

    def read_output(self) -> Output:
        phi = np.zeros(2 * self.ntargets)
        grad = np.zeros(6 * self.ntargets) if self.gradients else None
        _lib.check(_lib.lib().fmmh_helmholtz_read_output(self._h, _lib.ptr(phi), _lib.ptr(grad), self.ntargets, 0))
        return Output(phi.view(np.complex128),
                      grad.view(np.complex128).reshape(-1, 3) if self.gradients else None)

    def expansions(self, locals_: bool, level: int) -> np.ndarray:
        n = (self.p + 1) ** 2
        nb = len(self.plan.box_keys(level))
        out = np.zeros((nb, 2, n))
        _lib.check(_lib.lib().fmmh_helmholtz_read_expansions(self._h, int(locals_), int(level), _lib.ptr(out)))
        return out[:, 0] + 1j * out[:, 1]

    def kind_timing(self, enable=True):
        _lib.check(_lib.lib().fmmh_helmholtz_kind_timing(self._h, int(enable)))

    def kind_times(self):
        ms = np.zeros(len(KINDS))
        n = np.zeros(len(KINDS), np.int64)
        _lib.check(_lib.lib().fmmh_helmholtz_kind_times(self._h, _lib.ptr(ms), _lib.ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KINDS)}


class HelmholtzFmm:
    """A Helmholtz FMM built for fixed sources and targets (fmm-exec's HelmholtzFmm,
    fmm-exec/src/helmholtz/fmm.rs:2202-2300)."""

    def __init__(self, plan: Plan, operator: HelmholtzOperator):
        self.plan = plan
        self.operator = operator

    @property
    def p(self) -> int:
        return self.operator.p

    def evaluate(self, charges) -> Output:
        """Complex potential and gradient at the targets (1/(4 pi) included), caller order."""
        op = self.operator
        op.load_charges(charges)
        op.reset()
        op.evaluate()
        return op.read_output()


class HelmholtzFmmBuilder:
    """Settings of a Helmholtz FMM (fmm-exec/src/helmholtz/fmm.rs:419-1017): the wavenumber
    (Wavenumber::new: k != 0, Re k >= 0, Im k >= 0) and the degrees, a tolerance for the
    degree rule (Degrees::Tolerance) or one degree for every level."""

    def __init__(self, k, tolerance: Optional[float] = 1e-6, degree: Optional[int] = None):
        k = complex(k)
        if k == 0 or k.real < 0 or k.imag < 0 or not np.isfinite(k):
            raise FmmError(9, "the wavenumber must be nonzero with Re k >= 0 and Im k >= 0")
        if degree is not None:
            tolerance = None  # a named degree replaces the rule
        elif tolerance is None or not 0.0 < tolerance < 1.0:
            raise FmmError(9, "give a tolerance in (0, 1) or a degree")
        self._k = k
        self._tol = tolerance
        self._degree = degree
        self._gradients = True
        self._max_level = 16
        self._leaf_points = 256
        self._precision = "f64"
        self._device = 0
        self._near = None
        self._m2l = "dense"
        self._dense_budget = 0

    def gradients(self, on: bool):
        self._gradients = bool(on)
        return self

    def max_level(self, level: int):
        self._max_level = int(level)
        return self

    def max_points_per_leaf(self, n: int):
        self._leaf_points = int(n)
        return self

    def precision(self, precision: str):
        if precision not in PRECISIONS:
            raise FmmError(3, f"unknown precision {precision!r}")
        self._precision = precision
        return self

    def m2l(self, method: str):
        """The M2L: "dense" (the default), "rotation" (the reference's RCR per pair) or
        "plane_wave" (the exponential M2L by merge and shift where the reference's rule gives
        a level a quadrature, RCR elsewhere); HelmholtzM2l in
        fmm-exec/src/helmholtz/operator.rs:448-479."""
        if method not in ("dense", "rotation", "plane_wave"):
            raise FmmError(7, f"Helmholtz M2L: dense, rotation or plane_wave, not {method!r}")
        self._m2l = method
        return self

    def dense_budget(self, nbytes: int):
        """Bytes of dense M2L tables (316 x 2 (p+1)^4 f64 per level, from the finest level up;
        0 the reference's DEFAULT_DENSE_BUDGET of 1 GiB, negative no limit); the levels
        beyond it take the per-pair rotation."""
        self._dense_budget = int(nbytes)
        return self

    def device(self, ordinal: int):
        self._device = int(ordinal)
        return self

    def near_wx(self, rule):
        """W and X pairs to P2P (WxNear, fmm-exec/src/helmholtz/fmm.rs:370-400): None or
        "off"; (w, x) crossovers in points; or "device", this GPU path's cost model: an M2P
        (P2L) costs about (p+1)^2 pair interactions per point, so w = x = (p+1)^2 at the
        FMM's degree (resolved at build)."""
        if rule is None or rule == "off":
            self._near = None
        elif rule == "device":
            self._near = "device"
        else:
            self._near = (float(rule[0]), float(rule[1]))
        return self

    def degree_for(self, plan: Plan) -> int:
        """One degree for every level: the rule's largest over the levels that translate
        (2 and deeper; levels 0 and 1 get degree 0 under the rule, degree.rs:55-58)."""
        if self._degree is not None:
            return int(self._degree)
        _, side = plan.domain()
        degs = [helmholtz_degree(self._k, side / 2 ** (l + 1), self._tol) for l in range(2, max(plan.nlevels, 3))]
        return max(degs)

    def build(self, sources, targets=None) -> HelmholtzFmm:
        near = self._near
        if near == "device":
            # the degree depends only on the domain and the levels, which the near rule
            # leaves unchanged: resolve it on the plan without the rule
            p = self.degree_for(Plan(sources, targets, self._leaf_points, self._max_level, None))
            near = (float((p + 1) ** 2),) * 2
            plan = Plan(sources, targets, self._leaf_points, self._max_level, near)
        else:
            plan = Plan(sources, targets, self._leaf_points, self._max_level, near)
            p = self.degree_for(plan)
        op = HelmholtzOperator(plan, p, self._k, self._precision, self._gradients, self._device, self._m2l,
                               self._dense_budget)
        return HelmholtzFmm(plan, op)
