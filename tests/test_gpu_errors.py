"""Device errors of an evaluation are kept (fmm-exec/src/device.rs:414-425), as the
reference's injected-error test checks them (fmm-exec/tests/mpi_exec.rs:5176-5260, on one
rank): the evaluation with the injected error and the next one both fail with a device
error, the first saying "injected"; a new build evaluates again, bit for bit the first
output.  The hook sets a host flag only: no kernel faults."""
import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib
from oracle import oracle as O

DEVICE = 8  # FMMH_DEVICE
INVALID = 9  # FMMH_INVALID_ARGUMENT


def test_inject_needs_an_operator():
    assert _lib.lib().fmmh_inject_error(None) == INVALID


@pytest.fixture(scope="module")
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def _build(pts, precision):
    return fmm_amd.FmmBuilder(6).precision(precision).max_points_per_leaf(64).build(pts)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_injected_error_is_kept_and_a_new_build_evaluates(_gpu, precision):
    rng = O.Rng(0x7808)
    pts = rng.cube(12000)
    q = rng.normals(len(pts))
    fmm = _build(pts, precision)
    first = fmm.evaluate(q)
    # an argument error is returned and not kept
    qq = np.ones(len(pts) + 1)
    assert _lib.lib().fmmh_load_charges(fmm.operator._h, _lib.ptr(qq), len(qq), 0) == INVALID
    again = fmm.evaluate(q)
    assert np.array_equal(again.potential, first.potential)
    fmm.inject_device_error()
    for round_ in ("the evaluation with the error", "the next evaluation"):
        with pytest.raises(_lib.FmmError) as e:
            fmm.evaluate(q)
        assert e.value.code == DEVICE, round_
        assert "injected" in str(e.value), round_
    # every call that touches the device returns the kept error, level calls included
    for call in (lambda: fmm.operator.sync(), lambda: fmm.operator.call("p2m", fmm.plan.nlevels - 1),
                 lambda: fmm.operator.expansions(False, 2)):
        with pytest.raises(_lib.FmmError) as e:
            call()
        assert e.value.code == DEVICE
    # a new build evaluates again, bit for bit the first output
    out = _build(pts, precision).evaluate(q)
    assert np.array_equal(out.potential, first.potential)
    assert np.array_equal(out.gradient, first.gradient)
