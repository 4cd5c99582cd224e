"""ctypes binding of libfmm_amd.so (the C ABI of include/fmm_amd.h).

The library is built in-tree by ``fmm_amd.build()`` (hipcc for gfx950) and loaded from
this directory; there is no fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
# FMM_AMD_LIBRARY selects another build of the same sources (e.g. a tuning variant)
LIB_PATH = os.environ.get("FMM_AMD_LIBRARY") or os.path.join(HERE, "libfmm_amd.so")
CSRC = os.path.join(HERE, "csrc")

STATUS = {
    0: "Ok",
    1: "NotCompiled",
    2: "NoDevice",
    3: "UnsupportedPrecision",
    4: "OutOfMemory",
    5: "TooLarge",
    6: "WrongDevice",
    7: "UnsupportedLayout",
    8: "Device",
    9: "InvalidArgument",
}

# Every symbol include/fmm_amd.h declares.
EXPORTS = [
    "fmmh_last_error", "fmmh_abi_version", "fmmh_plan_build", "fmmh_plan_free", "fmmh_plan_nlevels",
    "fmmh_plan_nboxes", "fmmh_plan_nleaves", "fmmh_plan_domain", "fmmh_plan_box_keys", "fmmh_plan_leaf_keys",
    "fmmh_plan_list", "fmmh_plan_point_leaves", "fmmh_plan_leaf_coordinates", "fmmh_tables",
    "fmmh_plan_build_device", "fmmh_plan_leaf_cells", "fmmh_plan_box_cells", "fmmh_plan_point_starts", "fmmh_kernel_p2p", "fmmh_kernel_p2m", "fmmh_kernel_p2l",
    "fmmh_kernel_l2p", "fmmh_kernel_m2p", "fmmh_kernel_zero", "fmmh_kernel_gather_output",
    "fmmh_device_count", "fmmh_device_describe", "fmmh_operator_create", "fmmh_operator_free", "fmmh_operator_stats",
    "fmmh_load_charges", "fmmh_reset", "fmmh_p2m", "fmmh_m2m", "fmmh_m2l", "fmmh_p2l", "fmmh_l2l",
    "fmmh_l2p", "fmmh_m2p", "fmmh_p2p", "fmmh_evaluate", "fmmh_read_output", "fmmh_sync",
    "fmmh_kind_timing", "fmmh_kind_times", "fmmh_read_expansions", "fmmh_vjp", "fmmh_forest_build",
    "fmmh_evaluate_graph", "fmmh_direct_sum", "fmmh_m2l_rotation_table",
    "fmmh_helmholtz_degree", "fmmh_helmholtz_create", "fmmh_helmholtz_free", "fmmh_helmholtz_stats",
    "fmmh_helmholtz_load_charges", "fmmh_helmholtz_reset", "fmmh_helmholtz_evaluate",
    "fmmh_helmholtz_read_output", "fmmh_helmholtz_read_expansions", "fmmh_helmholtz_kind_timing",
    "fmmh_helmholtz_kind_times", "fmmh_helmholtz_table", "fmmh_helmholtz_direct_sum",
    "fmmh_plane_wave_quadrature", "fmmh_plane_wave_m2l", "fmmh_plane_wave_merge", "fmmh_plane_wave_slot_direction",
    "fmmh_helmholtz_plane_wave_quadrature", "fmmh_bessel_j_norms", "fmmh_helmholtz_plane_wave_m2l",
    "fmmh_helmholtz_plane_wave_merge",
    "fmmh_stream_wait", "fmmh_helmholtz_stream_wait",
    "fmmh_direct_sum_noalloc", "fmmh_operator_create_tuned", "fmmh_autotune", "fmmh_inject_error",
    "fmmh_set_table_cache", "fmmh_table_cache_stats",
]


class FmmError(RuntimeError):
    """A failed call of the C ABI; ``status`` names the KernelError-like code."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code
        self.status = STATUS.get(code, str(code))


class PlanOptions(C.Structure):
    _fields_ = [("leaf_points", C.c_int64), ("max_level", C.c_int32), ("near_w", C.c_double),
                ("near_x", C.c_double)]


class OperatorOptions(C.Structure):
    _fields_ = [("p", C.c_int32), ("precision", C.c_int32), ("gradients", C.c_int32), ("device", C.c_int32),
                ("m2l", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("p2p_uniform_min", C.c_int32), ("m2l_stage", C.c_int32), ("m2l_mode", C.c_int32 * 17),
                ("l2l_width", C.c_int32 * 17),
                ("m2m_form", C.c_int32 * 17)]


class HelmholtzOptions(C.Structure):
    _fields_ = [("p", C.c_int32), ("k_re", C.c_double), ("k_im", C.c_double), ("precision", C.c_int32),
                ("gradients", C.c_int32), ("device", C.c_int32), ("m2l", C.c_int32),
                ("m2l_dense_budget", C.c_int64)]


class P2pInputs(C.Structure):
    """fmmh_p2p_inputs (device pointers)."""
    _fields_ = [("leaf_first", C.c_uint32), ("count", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("near_off", "near_ent", "leaf_cell", "src_start", "tgt_start", "sx", "sy", "sz",
                                  "q", "tx", "ty", "tz")]


class SourceInputs(C.Structure):
    """fmmh_source_inputs (device pointers)."""
    _fields_ = [("rows", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("row_box", "row_cell", "row_off", "row_ent", "leaf_cell", "src_start", "sx", "sy",
                                  "sz", "q")]


class TargetInputs(C.Structure):
    """fmmh_target_inputs (device pointers)."""
    _fields_ = [("rows", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("row_leaf", "row_off", "row_box", "box_cell", "leaf_cell", "tgt_start", "tx", "ty",
                                  "tz")]


class Stats(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("nleaves", C.c_int64), ("nboxes", C.c_int64), ("p2p_pairs", C.c_uint64),
                ("m2l_pairs", C.c_uint64), ("m2l_slots", C.c_uint64), ("device_bytes", C.c_uint64)]


_lib = None


def build(jobs: int = 8) -> str:
    """Compile every HIP/C++ source into fmm_amd/libfmm_amd.so (hipcc, gfx950)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", CSRC], check=True)
    return LIB_PATH


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FmmError(1, f"{LIB_PATH} is not built (run fmm_amd.build())")
    # One HIP runtime per process: when PyTorch is installed, its runtime is loaded first
    # and the library binds to it (same soname), so torch tensors and the library share
    # one runtime whichever is used first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
    L.fmmh_last_error.restype = C.c_char_p
    L.fmmh_abi_version.restype = i32
    L.fmmh_plan_build.argtypes = [vp, i64, vp, i64, C.POINTER(PlanOptions), C.POINTER(vp)]
    L.fmmh_plan_build.restype = i32
    L.fmmh_plan_free.argtypes = [vp]
    L.fmmh_plan_free.restype = None
    L.fmmh_plan_nlevels.argtypes = [vp]
    L.fmmh_plan_nlevels.restype = i32
    L.fmmh_plan_nboxes.argtypes = [vp, i32]
    L.fmmh_plan_nboxes.restype = i64
    L.fmmh_plan_nleaves.argtypes = [vp]
    L.fmmh_plan_nleaves.restype = i64
    L.fmmh_plan_domain.argtypes = [vp, vp, vp]
    L.fmmh_plan_box_keys.argtypes = [vp, i32, vp]
    L.fmmh_plan_leaf_keys.argtypes = [vp, vp]
    L.fmmh_plan_list.argtypes = [vp, i32, i32, vp, vp, vp]
    L.fmmh_plan_list.restype = i64
    L.fmmh_plan_point_leaves.argtypes = [vp, i32, vp]
    L.fmmh_plan_leaf_coordinates.argtypes = [vp, i32, vp, vp]
    L.fmmh_plan_leaf_cells.argtypes = [vp, vp]
    L.fmmh_plan_build_device.argtypes = [i32, vp, i64, vp, i64, C.POINTER(PlanOptions), C.POINTER(vp)]
    L.fmmh_plan_box_cells.argtypes = [vp, i32, vp]
    L.fmmh_plan_point_starts.argtypes = [vp, i32, vp]
    L.fmmh_kernel_p2p.argtypes = [i32, C.POINTER(P2pInputs), i32, vp, vp, vp]
    L.fmmh_kernel_p2m.argtypes = [i32, i32, C.POINTER(SourceInputs), vp, vp]
    L.fmmh_kernel_p2l.argtypes = [i32, i32, C.POINTER(SourceInputs), vp, vp]
    L.fmmh_kernel_l2p.argtypes = [i32, i32, i32, C.POINTER(TargetInputs), vp, vp, vp, vp]
    L.fmmh_kernel_m2p.argtypes = [i32, i32, i32, C.POINTER(TargetInputs), vp, vp, vp, vp]
    L.fmmh_kernel_zero.argtypes = [i32, vp, C.c_int64, vp]
    L.fmmh_kernel_gather_output.argtypes = [i32, C.c_uint32, vp, vp, vp, vp, vp, i32, vp, vp, vp]
    L.fmmh_tables.argtypes = [i32, vp, vp, vp]
    L.fmmh_set_table_cache.argtypes = [C.c_char_p]
    L.fmmh_table_cache_stats.argtypes = [vp, vp]
    L.fmmh_m2l_rotation_table.argtypes = [i32, i32, vp]
    L.fmmh_device_count.argtypes = [C.POINTER(i32)]
    L.fmmh_device_describe.argtypes = [i32, C.c_char_p, i32]
    L.fmmh_operator_create.argtypes = [vp, C.POINTER(OperatorOptions), C.POINTER(vp)]
    L.fmmh_operator_create_tuned.argtypes = [vp, C.POINTER(OperatorOptions), C.POINTER(Tuning), C.POINTER(vp)]
    L.fmmh_autotune.argtypes = [vp, C.POINTER(OperatorOptions), dbl, C.POINTER(Tuning), vp]
    L.fmmh_inject_error.argtypes = [vp]
    L.fmmh_operator_free.argtypes = [vp]
    L.fmmh_operator_free.restype = None
    L.fmmh_operator_stats.argtypes = [vp, C.POINTER(Stats)]
    L.fmmh_load_charges.argtypes = [vp, vp, i64, i32]
    L.fmmh_reset.argtypes = [vp]
    for name in ("p2m", "m2m", "m2l", "p2l", "l2l", "l2p", "m2p", "p2p"):
        getattr(L, "fmmh_" + name).argtypes = [vp, i32]
    L.fmmh_evaluate.argtypes = [vp]
    L.fmmh_evaluate_graph.argtypes = [vp]
    L.fmmh_direct_sum.argtypes = [i32, vp, vp, i64, vp, i64, vp, vp, i32]
    L.fmmh_direct_sum_noalloc.argtypes = [i32, vp, vp, i64, vp, i64, vp, vp, vp]
    L.fmmh_read_output.argtypes = [vp, vp, vp, i64, i32]
    L.fmmh_stream_wait.argtypes = [vp, vp]
    L.fmmh_sync.argtypes = [vp]
    L.fmmh_kind_timing.argtypes = [vp, i32]
    L.fmmh_read_expansions.argtypes = [vp, i32, i32, vp]
    L.fmmh_kind_times.argtypes = [vp, vp, vp]
    L.fmmh_vjp.argtypes = [vp, vp, vp, vp, i64, i32, vp, vp]
    L.fmmh_forest_build.argtypes = [vp, vp, i32, C.POINTER(PlanOptions), C.POINTER(vp)]
    L.fmmh_plane_wave_quadrature.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp]
    L.fmmh_plane_wave_m2l.argtypes = [i32, i32, i32, i32, vp, vp]
    L.fmmh_plane_wave_merge.argtypes = [i32, i32, vp, vp, vp]
    L.fmmh_plane_wave_slot_direction.argtypes = [i32, i32]
    L.fmmh_helmholtz_plane_wave_quadrature.argtypes = [dbl, dbl, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.fmmh_bessel_j_norms.argtypes = [dbl, dbl, i32, vp]
    L.fmmh_helmholtz_plane_wave_m2l.argtypes = [i32, dbl, dbl, i32, i32, vp, vp]
    L.fmmh_helmholtz_plane_wave_merge.argtypes = [i32, dbl, dbl, vp, vp, vp]
    L.fmmh_helmholtz_degree.argtypes = [dbl, dbl, dbl]
    L.fmmh_helmholtz_create.argtypes = [vp, C.POINTER(HelmholtzOptions), C.POINTER(vp)]
    L.fmmh_helmholtz_free.argtypes = [vp]
    L.fmmh_helmholtz_free.restype = None
    L.fmmh_helmholtz_stats.argtypes = [vp, C.POINTER(Stats)]
    L.fmmh_helmholtz_load_charges.argtypes = [vp, vp, i64, i32]
    L.fmmh_helmholtz_reset.argtypes = [vp]
    L.fmmh_helmholtz_evaluate.argtypes = [vp]
    L.fmmh_helmholtz_read_output.argtypes = [vp, vp, vp, i64, i32]
    L.fmmh_helmholtz_stream_wait.argtypes = [vp, vp]
    L.fmmh_helmholtz_read_expansions.argtypes = [vp, i32, i32, vp]
    L.fmmh_helmholtz_kind_timing.argtypes = [vp, i32]
    L.fmmh_helmholtz_kind_times.argtypes = [vp, vp, vp]
    L.fmmh_helmholtz_table.argtypes = [i32, dbl, dbl, i32, i32, vp]
    L.fmmh_helmholtz_direct_sum.argtypes = [i32, dbl, dbl, vp, vp, i64, vp, i64, vp, vp, i32]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is C.c_int:  # default
            f.restype = i32
    _lib = L
    return L


def check(code: int) -> None:
    if code != 0:
        msg = lib().fmmh_last_error()
        raise FmmError(code, msg.decode() if msg else "")


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)
