"""ctypes view of the oracle library (test infrastructure only).

The oracle is the f64 CPU restatement of the reference FMM that the GPU path is checked
against.  Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u64 = C.POINTER(C.c_uint64)
        L.orc_next_u64.argtypes = [u64]
        L.orc_next_u64.restype = C.c_uint64
        L.orc_uniform.argtypes = [u64]
        L.orc_uniform.restype = C.c_double
        L.orc_normal.argtypes = [u64]
        L.orc_normal.restype = C.c_double
        for name in ("orc_cube", "orc_ball", "orc_sphere", "orc_plummer"):
            getattr(L, name).argtypes = [u64, C.c_int64, _f64p, C.c_double, _f64p]
        L.orc_gaussian_clusters.argtypes = [u64, C.c_int64, _f64p, C.c_int64, C.c_double, _f64p]
        L.orc_charges.argtypes = [u64, C.c_int64, _f64p]
        L.orc_normals.argtypes = [u64, C.c_int64, _f64p]
        L.orc_sample.argtypes = [u64, C.c_int64, C.c_int64, _i64p]
        pd = C.c_void_p
        L.orc_p2p.argtypes = [_f64p, _f64p, C.c_int64, _f64p, C.c_int64, _f64p, pd]
        L.orc_direct_sum.argtypes = [_f64p, _f64p, C.c_int64, _f64p, C.c_int64, _f64p, pd]
        L.orc_helmholtz_direct_sum.argtypes = [C.c_double, C.c_double, _f64p, _f64p, C.c_int64, _f64p, C.c_int64,
                                               _f64p, pd]
        L.orc_neumaier_sum.argtypes = [_f64p, C.c_int64]
        L.orc_neumaier_sum.restype = C.c_double
        L.orc_error_norms.argtypes = [_f64p, _f64p, C.c_int64, C.c_int, _f64p]
        for name in ("orc_regular", "orc_irregular"):
            getattr(L, name).argtypes = [C.c_int, _f64p, _f64p]
        for name in ("orc_regular_grad", "orc_irregular_grad"):
            getattr(L, name).argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        for name in ("orc_p2m", "orc_p2l"):
            getattr(L, name).argtypes = [C.c_int, _f64p, C.c_double, _f64p, _f64p, C.c_int64, _f64p]
        for name in ("orc_l2p", "orc_m2p"):
            getattr(L, name).argtypes = [C.c_int, _f64p, C.c_double, _f64p, _f64p, C.c_int64, _f64p, pd]
        for name in ("orc_m2m", "orc_l2l", "orc_m2l"):
            getattr(L, name).argtypes = [C.c_int, _f64p, C.c_double, _f64p, _f64p]
        L.orc_offset_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_offset_index.restype = C.c_int
        L.orc_offset_of_index.argtypes = [C.c_int, C.POINTER(C.c_int)]
        for name in ("orc_table_m2m", "orc_table_l2l", "orc_table_m2l"):
            getattr(L, name).argtypes = [C.c_int, _f64p]
        L.orc_key.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64]
        L.orc_key.restype = C.c_uint64
        L.orc_decode.argtypes = [C.c_uint64, _i64p]
        L.orc_parent.argtypes = [C.c_uint64]
        L.orc_parent.restype = C.c_uint64
        L.orc_adjacent.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_adjacent.restype = C.c_int
        L.orc_bounding_box.argtypes = [_f64p, C.c_int64, _f64p]
        L.orc_point_key.argtypes = [_f64p, _f64p, C.c_int]
        L.orc_point_key.restype = C.c_uint64
        L.orc_fmm_new.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, C.c_int, C.c_int64, C.c_int, C.c_int,
                                  C.c_double, C.c_double]
        L.orc_fmm_new.restype = C.c_void_p
        L.orc_fmm_free.argtypes = [C.c_void_p]
        L.orc_fmm_evaluate.argtypes = [C.c_void_p, _f64p, _f64p, pd]
        L.orc_fmm_evaluate_cpu.argtypes = [C.c_void_p, _f64p, _f64p, pd, C.c_int]
        L.orc_cpu_isa.restype = C.c_int
        L.orc_cpu_threads.restype = C.c_int
        L.orc_fmm_nlevels.argtypes = [C.c_void_p]
        L.orc_fmm_nlevels.restype = C.c_int
        L.orc_fmm_domain.argtypes = [C.c_void_p, _f64p]
        L.orc_fmm_domain.restype = C.c_double
        L.orc_fmm_nboxes.argtypes = [C.c_void_p, C.c_int]
        L.orc_fmm_nboxes.restype = C.c_int64
        L.orc_fmm_box_keys.argtypes = [C.c_void_p, C.c_int, _u64p]
        L.orc_fmm_nleaves.argtypes = [C.c_void_p]
        L.orc_fmm_nleaves.restype = C.c_int64
        L.orc_fmm_leaf_keys.argtypes = [C.c_void_p, _u64p]
        L.orc_fmm_list.argtypes = [C.c_void_p, C.c_int, C.c_int, pd, pd, pd]
        L.orc_fmm_list.restype = C.c_int64
        L.orc_fmm_point_leaves.argtypes = [C.c_void_p, C.c_int, _i64p]
        L.orc_fmm_leaf_coordinates.argtypes = [C.c_void_p, C.c_int, _f64p, _i64p]
        L.orc_fmm_expansions.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p]
        L.orc_h_solid.argtypes = [C.c_int, _f64p, _f64p]
        L.orc_h_solid_grad.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.orc_h_radial.argtypes = [C.c_int, C.c_double, C.c_double, _f64p, _f64p]
        for name in ("orc_h_p2m", "orc_h_p2l"):
            getattr(L, name).argtypes = [C.c_int, C.c_double, C.c_double, _f64p, C.c_double, _f64p, _f64p,
                                         C.c_int64, _f64p]
        for name in ("orc_h_l2p", "orc_h_m2p"):
            getattr(L, name).argtypes = [C.c_int, C.c_double, C.c_double, _f64p, C.c_double, _f64p, _f64p,
                                         C.c_int64, _f64p, pd]
        L.orc_h_rotation_blocks.argtypes = [C.c_int, _f64p, _f64p]
        L.orc_h_coaxial.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _f64p]
        for name in ("orc_h_m2m", "orc_h_l2l"):
            getattr(L, name).argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _f64p]
        L.orc_h_m2l.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int), _f64p]
        L.orc_h_apply.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_hfmm_evaluate.argtypes = [C.c_void_p, C.c_double, C.c_double, _f64p, _f64p, pd]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Rng:
    """SplitMix64 (fmm-validate/src/rng.rs) with the point distributions."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed)

    def next_u64(self) -> int:
        return lib().orc_next_u64(C.byref(self.state))

    def uniform(self) -> float:
        return lib().orc_uniform(C.byref(self.state))

    def _pts(self, name, n, centre, scale):
        out = np.empty((n, 3))
        getattr(lib(), name)(C.byref(self.state), n, np.ascontiguousarray(centre, np.float64), scale, out)
        return out

    def cube(self, n, centre=(0.5, 0.5, 0.5), half=0.5):
        return self._pts("orc_cube", n, centre, half)

    def ball(self, n, centre, radius):
        return self._pts("orc_ball", n, centre, radius)

    def sphere(self, n, centre, radius):
        return self._pts("orc_sphere", n, centre, radius)

    def plummer(self, n, centre, scale):
        return self._pts("orc_plummer", n, centre, scale)

    def gaussian_clusters(self, n, centres, width):
        c = np.ascontiguousarray(centres, np.float64).reshape(-1, 3)
        out = np.empty((n, 3))
        lib().orc_gaussian_clusters(C.byref(self.state), n, c, len(c), width, out)
        return out

    def charges(self, n):
        out = np.empty(n)
        lib().orc_charges(C.byref(self.state), n, out)
        return out

    def normals(self, n):
        out = np.empty(n)
        lib().orc_normals(C.byref(self.state), n, out)
        return out

    def sample(self, n, sampled):
        out = np.empty(n, np.int64)
        lib().orc_sample(C.byref(self.state), n, sampled, out)
        return out[: min(sampled, n)].copy()


BENCH_SEED = 0x00BE9C06


def bench_problem(n: int, sampled: int, charges: str = "uniform"):
    """The problem of fmm-bench/src/measure.rs:44-67: cube points, then charges
    (uniform[-1, 1), or N(0, 1) by Box-Muller), then the Fisher-Yates sample."""
    rng = Rng(BENCH_SEED)
    pts = rng.cube(n)
    q = rng.charges(n) if charges == "uniform" else rng.normals(n)
    sample = rng.sample(n, sampled)
    return pts, q, sample


def cpu_isa() -> str:
    """The kernels the CPU baseline runs here: "avx512f" or "scalar"."""
    return "avx512f" if lib().orc_cpu_isa() == 2 else "scalar"


def cpu_threads() -> int:
    return int(lib().orc_cpu_threads())


def direct_sum(src, q, tgt, gradients=True):
    src = np.ascontiguousarray(src, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    phi = np.zeros(len(tgt))
    grad = np.zeros((len(tgt), 3)) if gradients else None
    lib().orc_direct_sum(src, q, len(src), tgt, len(tgt), phi, _ptr(grad))
    return phi, grad


def helmholtz_direct_sum(k, src, q, tgt, gradients=True):
    """Compensated f64 direct sum of q e^{ikr}/r and its gradient (complex k and q; no
    1/(4 pi)); returns complex arrays (nt,) and (nt, 3)."""
    src = np.ascontiguousarray(src, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    qc = np.ascontiguousarray(np.asarray(q, np.complex128).view(np.float64))
    pot = np.zeros(2 * len(tgt))
    grad = np.zeros(6 * len(tgt)) if gradients else None
    lib().orc_helmholtz_direct_sum(float(np.real(k)), float(np.imag(k)), src, qc, len(src), tgt, len(tgt), pot,
                                   _ptr(grad))
    g = grad.view(np.complex128).reshape(-1, 3) if gradients else None
    return pot.view(np.complex128), g


def h_solid(p, x):
    out = np.zeros(nc(p))
    lib().orc_h_solid(p, np.ascontiguousarray(x, np.float64), out)
    return out


def h_solid_grad(p, x):
    g = np.zeros((3, nc(p)))
    lib().orc_h_solid_grad(p, np.ascontiguousarray(x, np.float64), g[0], g[1], g[2])
    return g


def h_radial(P, z):
    """Spherical j_n(z) and h_n^(1)(z), n = 0..P, for complex z."""
    j, h = np.zeros(2 * (P + 1)), np.zeros(2 * (P + 1))
    lib().orc_h_radial(P, float(np.real(z)), float(np.imag(z)), j, h)
    return j.view(np.complex128), h.view(np.complex128)


def _cplx_planes(x):
    x = np.asarray(x, np.complex128)
    return np.ascontiguousarray(np.concatenate([x.real, x.imag]))


def _from_planes(v):
    n = len(v) // 2
    return v[:n] + 1j * v[n:]


def h_p2m(p, k, c, r, src, q):
    out = np.zeros(2 * nc(p))
    qc = np.ascontiguousarray(np.asarray(q, np.complex128).view(np.float64))
    lib().orc_h_p2m(p, float(np.real(k)), float(np.imag(k)), np.asarray(c, np.float64), float(r),
                    np.ascontiguousarray(src, np.float64), qc, len(src), out)
    return _from_planes(out)


def h_p2l(p, k, c, r, src, q):
    out = np.zeros(2 * nc(p))
    qc = np.ascontiguousarray(np.asarray(q, np.complex128).view(np.float64))
    lib().orc_h_p2l(p, float(np.real(k)), float(np.imag(k)), np.asarray(c, np.float64), float(r),
                    np.ascontiguousarray(src, np.float64), qc, len(src), out)
    return _from_planes(out)


def _h_eval(name, p, k, c, r, coeffs, tgt, gradients):
    tgt = np.ascontiguousarray(tgt, np.float64)
    phi = np.zeros(2 * len(tgt))
    grad = np.zeros(6 * len(tgt)) if gradients else None
    getattr(lib(), name)(p, float(np.real(k)), float(np.imag(k)), np.asarray(c, np.float64), float(r),
                         _cplx_planes(coeffs), tgt, len(tgt), phi, _ptr(grad))
    return phi.view(np.complex128), (grad.view(np.complex128).reshape(-1, 3) if gradients else None)


def h_l2p(p, k, c, r, local, tgt, gradients=True):
    return _h_eval("orc_h_l2p", p, k, c, r, local, tgt, gradients)


def h_m2p(p, k, c, r, mult, tgt, gradients=True):
    return _h_eval("orc_h_m2p", p, k, c, r, mult, tgt, gradients)


def h_rotation_blocks(p, Q):
    total = sum((2 * n + 1) ** 2 for n in range(p + 1))
    out = np.zeros(total)
    lib().orc_h_rotation_blocks(p, np.ascontiguousarray(Q, np.float64).ravel(), out)
    blocks, off = [], 0
    for n in range(p + 1):
        w = 2 * n + 1
        blocks.append(out[off:off + w * w].reshape(w, w))
        off += w * w
    return blocks


def h_coaxial(p, singular_start, k, t):
    """E^m_{n'n}(t) as an array [m, n', n] (complex)."""
    out = np.zeros(2 * (p + 1) ** 3)
    lib().orc_h_coaxial(p, int(singular_start), float(np.real(k)), float(np.imag(k)), float(t), out)
    return out.view(np.complex128).reshape(p + 1, p + 1, p + 1)


def _h_table(out, p):
    n = nc(p)
    return (out[: n * n] + 1j * out[n * n:]).reshape(n, n).T  # row = target slot


def h_m2m(p, k, o):
    out = np.zeros(2 * nc(p) ** 2)
    lib().orc_h_m2m(p, float(np.real(k)), float(np.imag(k)), int(o), out)
    return _h_table(out, p)


def h_l2l(p, k, o):
    out = np.zeros(2 * nc(p) ** 2)
    lib().orc_h_l2l(p, float(np.real(k)), float(np.imag(k)), int(o), out)
    return _h_table(out, p)


def h_m2l(p, k, d):
    out = np.zeros(2 * nc(p) ** 2)
    dv = (C.c_int * 3)(*[int(v) for v in d])
    lib().orc_h_m2l(p, float(np.real(k)), float(np.imag(k)), dv, out)
    return _h_table(out, p)


def p2p(src, q, tgt, gradients=True):
    src = np.ascontiguousarray(src, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    phi = np.zeros(len(tgt))
    grad = np.zeros((len(tgt), 3)) if gradients else None
    lib().orc_p2p(src, q, len(src), tgt, len(tgt), phi, _ptr(grad))
    return phi, grad


def error_norms(approx, exact, vec3=False):
    out = np.zeros(2)
    a = np.ascontiguousarray(approx, np.float64)
    e = np.ascontiguousarray(exact, np.float64)
    lib().orc_error_norms(a, e, len(e) if not vec3 else len(e.reshape(-1, 3)), int(vec3), out)
    return float(out[0]), float(out[1])


def nc(p):
    return (p + 1) * (p + 1)


def regular(p, x):
    out = np.zeros(nc(p))
    lib().orc_regular(p, np.asarray(x, np.float64), out)
    return out


def irregular(p, x):
    out = np.zeros(nc(p))
    lib().orc_irregular(p, np.asarray(x, np.float64), out)
    return out


def regular_grad(p, x):
    g = np.zeros((3, nc(p)))
    lib().orc_regular_grad(p, np.asarray(x, np.float64), g[0], g[1], g[2])
    return g


def irregular_grad(p, x):
    g = np.zeros((3, nc(p)))
    lib().orc_irregular_grad(p, np.asarray(x, np.float64), g[0], g[1], g[2])
    return g


def p2m(p, c, r, src, q):
    out = np.zeros(nc(p))
    src = np.ascontiguousarray(src, np.float64)
    lib().orc_p2m(p, np.asarray(c, np.float64), r, src, np.ascontiguousarray(q, np.float64), len(src), out)
    return out


def p2l(p, c, r, src, q):
    out = np.zeros(nc(p))
    src = np.ascontiguousarray(src, np.float64)
    lib().orc_p2l(p, np.asarray(c, np.float64), r, src, np.ascontiguousarray(q, np.float64), len(src), out)
    return out


def _eval(name, p, c, r, coeffs, tgt, gradients):
    tgt = np.ascontiguousarray(tgt, np.float64)
    phi = np.zeros(len(tgt))
    grad = np.zeros((len(tgt), 3)) if gradients else None
    getattr(lib(), name)(p, np.asarray(c, np.float64), r, np.ascontiguousarray(coeffs, np.float64),
                         tgt, len(tgt), phi, _ptr(grad))
    return phi, grad


def l2p(p, c, r, local, tgt, gradients=True):
    return _eval("orc_l2p", p, c, r, local, tgt, gradients)


def m2p(p, c, r, mult, tgt, gradients=True):
    return _eval("orc_m2p", p, c, r, mult, tgt, gradients)


def _translate(name, p, s, ratio, x):
    out = np.zeros(nc(p))
    getattr(lib(), name)(p, np.asarray(s, np.float64), ratio, np.ascontiguousarray(x, np.float64), out)
    return out


def m2m(p, b, rho, x):
    return _translate("orc_m2m", p, b, rho, x)


def l2l(p, t, sigma, x):
    return _translate("orc_l2l", p, t, sigma, x)


def m2l(p, b, sigma, x):
    return _translate("orc_m2l", p, b, sigma, x)


def offset_index(d):
    return lib().orc_offset_index(int(d[0]), int(d[1]), int(d[2]))


def offset_of_index(i):
    d = (C.c_int * 3)()
    lib().orc_offset_of_index(i, d)
    return tuple(d)


def tables(p):
    """(M2M[8], L2L[8], M2L[316]) as arrays of (p+1)^2 x (p+1)^2 matrices A with
    output = A @ input (stored column-major in the library)."""
    n = nc(p)
    out = {}
    for name, count in (("m2m", 8), ("l2l", 8), ("m2l", 316)):
        buf = np.zeros(count * n * n)
        getattr(lib(), "orc_table_" + name)(p, buf)
        out[name] = buf.reshape(count, n, n).transpose(0, 2, 1).copy()
    return out["m2m"], out["l2l"], out["m2l"]


def key(level, i, j, k):
    return lib().orc_key(level, i, j, k)


def decode(key_):
    out = np.zeros(4, np.int64)
    lib().orc_decode(key_, out)
    return tuple(int(v) for v in out)


def adjacent(a, b):
    return bool(lib().orc_adjacent(a, b))


class Fmm:
    """The oracle FMM (one rank, near rule off): tree, lists and f64 evaluation."""

    def __init__(self, sources, targets=None, p=6, leaf_points=64, max_level=16, gradients=True, near_rule=None):
        self.sources = np.ascontiguousarray(sources, np.float64)
        self.targets = self.sources if targets is None else np.ascontiguousarray(targets, np.float64)
        self.gradients = gradients
        self.p = p
        nw, nx = (-1.0, -1.0) if near_rule is None else (float(near_rule[0]), float(near_rule[1]))
        self._h = lib().orc_fmm_new(self.sources, len(self.sources), self.targets, len(self.targets),
                                   p, leaf_points, max_level, int(gradients), nw, nx)
        if not self._h:
            raise ValueError("degenerate point set")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_fmm_free(self._h)
            self._h = None

    def evaluate(self, charges):
        q = np.ascontiguousarray(charges, np.float64)
        phi = np.zeros(len(self.targets))
        grad = np.zeros((len(self.targets), 3)) if self.gradients else None
        lib().orc_fmm_evaluate(self._h, q, phi, _ptr(grad))
        return phi, grad

    def evaluate_cpu_path(self, charges, threads=0):
        """The same FMM by the CPU baseline (cpu_path.c): the reference's multi-threaded
        host evaluation restated (threads over targets, AVX-512 P2P, batched M2L column
        kernel with fused steps); threads <= 0 uses OpenMP's default."""
        q = np.ascontiguousarray(charges, np.float64)
        phi = np.zeros(len(self.targets))
        grad = np.zeros((len(self.targets), 3)) if self.gradients else None
        lib().orc_fmm_evaluate_cpu(self._h, q, phi, _ptr(grad), int(threads))
        return phi, grad

    def evaluate_helmholtz(self, k, charges):
        """Helmholtz FMM on the same tree (wavenumber k, complex charges): complex phi (nt,)
        and grad (nt, 3)."""
        q = np.ascontiguousarray(np.asarray(charges, np.complex128).view(np.float64))
        phi = np.zeros(2 * len(self.targets))
        grad = np.zeros(6 * len(self.targets)) if self.gradients else None
        lib().orc_hfmm_evaluate(self._h, float(np.real(k)), float(np.imag(k)), q, phi, _ptr(grad))
        g = grad.view(np.complex128).reshape(-1, 3) if self.gradients else None
        return phi.view(np.complex128), g

    @property
    def nlevels(self):
        return lib().orc_fmm_nlevels(self._h)

    def domain(self):
        box = np.zeros(6)
        side = lib().orc_fmm_domain(self._h, box)
        return box, side

    def box_keys(self, level):
        n = lib().orc_fmm_nboxes(self._h, level)
        out = np.zeros(n, np.uint64)
        if n:
            lib().orc_fmm_box_keys(self._h, level, out)
        return out

    def leaf_keys(self):
        n = lib().orc_fmm_nleaves(self._h)
        out = np.zeros(n, np.uint64)
        lib().orc_fmm_leaf_keys(self._h, out)
        return out

    def list(self, kind: str, level: int, aux: bool = False):
        """CSR (offsets, entries[, aux]) of 'near', 'v', 'w' or 'x' rows; aux (always for
        V, on request for W): the offset index of each V entry, the level of each W
        entry's box."""
        k = {"near": 0, "v": 1, "w": 2, "x": 3}[kind]
        total = lib().orc_fmm_list(self._h, k, level, None, None, None)
        if kind in ("near", "w"):
            lk = self.leaf_keys()
            levels = lk & np.uint64(0x7FFF)
            rows = int(np.sum(levels == level))
        else:
            rows = len(self.box_keys(level))
        off = np.zeros(rows + 1, np.int64)
        ent = np.zeros(max(total, 1), np.int64)
        ax = np.zeros(max(total, 1), np.int64)
        lib().orc_fmm_list(self._h, k, level, _ptr(off), _ptr(ent), _ptr(ax))
        if kind == "v" or (kind == "w" and aux):
            return off, ent[:total], ax[:total]
        return off, ent[:total]

    def leaf_coordinates(self, targets=False):
        n = len(self.targets) if targets else len(self.sources)
        u, order = np.zeros((n, 3)), np.zeros(n, np.int64)
        lib().orc_fmm_leaf_coordinates(self._h, int(targets), u, order)
        return u, order

    def expansions(self, locals_: bool, level: int):
        out = np.zeros((len(self.box_keys(level)), nc(self.p)))
        if lib().orc_fmm_expansions(self._h, int(locals_), level, out) != 0:
            raise ValueError("no evaluation yet")
        return out

    def point_leaves(self, targets=False):
        n = len(self.targets) if targets else len(self.sources)
        out = np.zeros(n, np.int64)
        lib().orc_fmm_point_leaves(self._h, int(targets), out)
        return out
