"""The autotuner of the launch shapes (fmm_amd/csrc/tune.cpp) and its cache file
(fmm_amd/tune.py), as the reference's device autotune (fmm-exec/src/tune.rs:1-110).

CPU: buckets, the cache's round trip, version and NaN rules, the plan's buckets, and the
C ABI's refusal of invalid decisions (before any device work).  GPU: the tuner's
decisions are valid, a second build takes them from the cache without timing anything,
and operators built with every decision forced (each mode, width and form) still match
the oracle: f64 to 1e-12, mixed to 1e-6.
"""
import ctypes as C
import json
import math

import numpy as np
import pytest

import fmm_amd
from fmm_amd import _lib, tune
from oracle import oracle as O


def test_bucket_rounds_up_to_powers_of_two():
    assert [tune.bucket(n) for n in (0, 1, 2, 3, 4, 5, 1000, 1024, 1025)] == [0, 1, 2, 4, 4, 8, 1024, 1024, 2048]
    assert tune.bucket(17.2) == 32


def test_cache_round_trip(tmp_path):
    buckets = {"p2p": {0: 64}, "m2l": {2: 256, 3: 16384}, "l2l": {1: 8, 2: 64, 3: 512}, "m2m": {0: 8, 1: 64, 2: 512},
               "stage": {0: 32768}}
    t = _lib.Tuning()
    t.p2p_uniform_min = 49
    t.m2l_stage = 2
    t.m2l_mode[2], t.m2l_mode[3] = 2, 1
    t.l2l_width[3] = 64
    t.m2m_form[0] = 1
    report = np.zeros((5, 17, 5))
    report[2, 1, 0] = math.nan  # L2L of level 1 not timed: not stored
    c = tune.TuningCache(tmp_path)
    c.store("k", buckets, t, report, "mixed")
    c.save()
    doc = json.load(open(tmp_path / tune.CACHE_NAME))
    assert doc["version"] == tune.CANDIDATE_SET_VERSION
    c2 = tune.TuningCache(tmp_path)
    t2, missing = c2.lookup("k", buckets, "mixed")
    assert missing == [("l2l", 1)]
    assert t2.p2p_uniform_min == 49 and t2.m2l_stage == 2 and t2.m2l_mode[2] == 2 and t2.m2l_mode[3] == 1
    assert t2.l2l_width[3] == 64 and t2.m2m_form[0] == 1
    # another key, or another candidate-set version, finds nothing
    assert len(c2.lookup("other", buckets, "mixed")[1]) == 10
    assert len(c2.lookup("other", buckets, "f64")[1]) == 8  # no P2P form or M2L stage in f64
    doc["version"] = -1
    json.dump(doc, open(tmp_path / tune.CACHE_NAME, "w"))
    assert tune.TuningCache(tmp_path).entries == {}


def test_level_buckets_of_a_plan():
    rng = O.Rng(81)
    pts = rng.cube(20000)
    plan = fmm_amd.Plan(pts, leaf_points=32)
    b = tune.level_buckets(plan)
    for l, v in b["m2l"].items():
        assert v == tune.bucket(len(plan.list("v", l)[1]))
    assert b["p2p"][0] == tune.bucket(20000 / plan.nleaves)
    assert set(b["l2l"]) == set(range(1, plan.nlevels))


def test_device_key_names_the_architecture():
    # the cache key's device part comes from the HIP runtime, never from a name that a
    # profiler or another binding may report differently
    buf = C.create_string_buffer(8)
    if fmm_amd.device_count() < 1:
        assert _lib.lib().fmmh_device_describe(0, buf, len(buf)) == 9
        return
    assert _lib.lib().fmmh_device_describe(0, buf, 2) == 9  # too small for any name
    name = tune.device_name(0)
    assert name.startswith("gfx") and name.endswith("cu")
    assert tune.cache_key(8, "mixed", True, "dense").startswith(name + "|p8|")


def test_invalid_tuning_refused():
    rng = O.Rng(82)
    plan = fmm_amd.Plan(rng.cube(500), leaf_points=32)
    o = _lib.OperatorOptions(4, 0, 1, 0, 0)
    h = C.c_void_p()
    for field, value in (("l2l_width", 32), ("m2l_mode", 3), ("m2m_form", -1), ("m2l_stage", 3)):
        t = _lib.Tuning()
        if field == "m2l_stage":
            t.m2l_stage = value
        else:
            getattr(t, field)[2] = value
        code = _lib.lib().fmmh_operator_create_tuned(plan.handle, C.byref(o), C.byref(t), C.byref(h))
        assert code == 9, field  # InvalidArgument, before any device work


@pytest.fixture(scope="module")
def _gpu():
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def _case():
    rng = O.Rng(83)
    pts = np.vstack([rng.cube(30000), rng.plummer(10000, (0.5, 0.5, 0.5), 0.05)])
    q = rng.normals(len(pts))
    return pts, q


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_autotune_and_cache(_gpu, tmp_path, precision):
    pts, q = _case()
    b = fmm_amd.FmmBuilder(8).precision(precision).max_points_per_leaf(64).tuning_cache(tmp_path).tuning_budget(20)
    f1 = b.build(pts)
    info = b.tuning_info
    print(f"autotune {precision}: {info['decisions']} in {info['tuned_s']:.1f} s", flush=True)
    assert not info["cached"] and info["tuned_s"] > 0
    d = info["decisions"]
    if precision == "mixed":
        assert d["p2p_uniform_min"] in (33, 17, 49, 65, 129)
    assert all(m in (0, 1, 2) for m in d["m2l_mode"])
    assert all(w in (0, 16, 64, 128) for w in d["l2l_width"])
    # the second build takes every decision from the cache
    b2 = fmm_amd.FmmBuilder(8).precision(precision).max_points_per_leaf(64).tuning_cache(tmp_path)
    f2 = b2.build(pts)
    assert b2.tuning_info["cached"] and b2.tuning_info["tuned_s"] == 0.0
    assert b2.tuning_info["decisions"] == d
    ref = O.Fmm(pts, p=8, leaf_points=64)
    rphi, rgrad = ref.evaluate(q)
    bar = 1e-12 if precision == "f64" else 1e-6
    for f in (f1, f2):
        out = f.evaluate(q)
        assert O.error_norms(out.potential, rphi)[0] < bar
        assert O.error_norms(out.gradient, rgrad, True)[0] < bar


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_every_forced_decision_matches_oracle(_gpu, precision):
    # each candidate of every decision, forced on every level, against the oracle
    pts, q = _case()
    plan = fmm_amd.Plan(pts, leaf_points=64)
    ref = O.Fmm(pts, p=8, leaf_points=64)
    rphi, rgrad = ref.evaluate(q)
    bar = 1e-12 if precision == "f64" else 1e-6
    variants = []
    for mode in (1, 2):
        for width, form, umin, stage in ((16, 1, 17, 2), (64, 2, 65, 1), (128, 1, 129, 2)):
            t = _lib.Tuning()
            t.p2p_uniform_min = umin
            t.m2l_stage = stage
            for l in range(17):
                t.m2l_mode[l] = mode
                t.l2l_width[l] = width
                t.m2m_form[l] = form
            variants.append((mode, width, form, umin, stage, t))
    for mode, width, form, umin, stage, t in variants:
        op = fmm_amd.Operator(plan, 8, precision, True, 0, "dense", t)
        f = fmm_amd.Fmm(plan, op)
        out = f.evaluate(q)
        e1 = O.error_norms(out.potential, rphi)[0]
        e2 = O.error_norms(out.gradient, rgrad, True)[0]
        print(f"forced {precision} m2l {mode} l2l {width} m2m {form} p2p {umin} stage {stage}: {e1:.2e} {e2:.2e}",
              flush=True)
        assert e1 < bar and e2 < bar


@pytest.mark.gpu
@pytest.mark.parametrize("p", [6, 10, 12])
def test_m2l_stages_match_oracle(_gpu, p):
    # the bf16x3 contraction in both stage shapes (one k-step per stage in a ring of three;
    # two in a ring of two, even k-step counts: p = 10 has 4, p = 12 has 6; p = 6 always
    # takes its two k-steps per stage) against the oracle
    rng = O.Rng(84)
    pts = rng.cube(60000)
    q = rng.normals(len(pts))
    plan = fmm_amd.Plan(pts, leaf_points=64)
    ref = O.Fmm(pts, p=p, leaf_points=64)
    rphi, rgrad = ref.evaluate(q)
    outs = []
    for stage in (0, 2):
        t = _lib.Tuning()
        t.m2l_stage = stage
        out = fmm_amd.Fmm(plan, fmm_amd.Operator(plan, p, "mixed", True, 0, "dense", t)).evaluate(q)
        e1 = O.error_norms(out.potential, rphi)[0]
        e2 = O.error_norms(out.gradient, rgrad, True)[0]
        print(f"m2l stage {stage} p={p}: vs oracle {e1:.2e} {e2:.2e}", flush=True)
        assert e1 < 1e-6 and e2 < 1e-6
        outs.append(out)
    # the same products and sums in the same order: the same bits
    assert np.array_equal(outs[0].potential, outs[1].potential)
