"""Autotune of the device operator's launch shapes, with an on-disk cache of decisions.

The role of the reference's device autotune and its tuning cache (fmm-exec/src/tune.rs:1-110,
FmmBuilder::tuning_cache): decisions are timed on the FMM's own levels at build
(fmmh_autotune, fmm_amd/csrc/tune.cpp), kept per bucket in a JSON file in a directory
the caller names, and reused by later builds whose levels fall in the same buckets.  A
bucket is a count rounded up to a power of two: the level's V pairs (M2L), the level's
boxes (L2L), the child level's boxes (M2M into the level), and the mean points per leaf
(P2P), as in the reference.  Where the cache has no decision and no budget is left, the
static rule applies (0 in the tuning).

Decisions (the static rule's candidate is timed first and wins ties):
  p2p  the mixed P2P's uniform-source threshold: 33 (static), 17, 49, 65, 129 (never)
  m2l  per level: box mode (1) or clustered mode (2)
  l2l  per level: 16, 64 or 128 columns per task
  m2m  per parent level: per-tile form (1) or split form (2)
  stage the mixed dense M2L's pipeline: one k-step per stage (0) or two in a ring of two
       buffers (2), by the bucket of all the levels' V pairs
  leaf optional (tune_leaf_size): the max points per leaf among the caller's candidates,
       by whole evaluations on the caller's points
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import tempfile
import time

import numpy as np

from . import _lib

CACHE_NAME = "fmm_amd_tuning.json"
# bump when a candidate set or a decision's meaning changes: older files are ignored
CANDIDATE_SET_VERSION = 1
KINDS = ("p2p", "m2l", "l2l", "m2m", "stage")
P2P_CANDIDATES = (33, 17, 49, 65, 129)


def bucket(n: float) -> int:
    """n rounded up to a power of two (0 stays 0)."""
    n = int(math.ceil(n))
    return 0 if n <= 0 else 1 << (n - 1).bit_length()


def device_name(device: int = 0) -> str:
    """The HIP runtime's architecture name and CU count of `device` (the cache key's
    device part; the same under a profiler as without one)."""
    buf = C.create_string_buffer(256)
    _lib.check(_lib.lib().fmmh_device_describe(int(device), buf, len(buf)))
    return buf.value.decode().replace("|", "/")


def cache_key(p, precision, gradients, m2l, device=0) -> str:
    return f"{device_name(device)}|p{int(p)}|{precision}|g{int(bool(gradients))}|{m2l}"


def level_buckets(plan) -> dict:
    """The bucket of every decision of this plan: {kind: {level: bucket}}."""
    L = plan.nlevels
    boxes = [len(plan.box_keys(l)) for l in range(L)]
    out = {"m2l": {}, "l2l": {}, "m2m": {}, "p2p": {}, "stage": {}}
    total = 0
    for l in range(2, L):
        npairs = len(plan.list("v", l)[1])
        total += npairs
        if npairs:
            out["m2l"][l] = bucket(npairs)
    if total:
        out["stage"][0] = bucket(total)
    for l in range(1, L):
        out["l2l"][l] = bucket(boxes[l])
    for l in range(L - 1):
        out["m2m"][l] = bucket(boxes[l + 1])
    per_leaf = (len(plan.sources) + len(plan.targets)) / (2.0 * max(plan.nleaves, 1))
    out["p2p"][0] = bucket(per_leaf)
    return out


class TuningCache:
    """Decisions per key and bucket in <directory>/fmm_amd_tuning.json."""

    def __init__(self, directory):
        self.directory = os.fspath(directory)
        self.path = os.path.join(self.directory, CACHE_NAME)
        self.entries = {}
        if os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    doc = json.load(f)
                if doc.get("version") == CANDIDATE_SET_VERSION:
                    self.entries = doc.get("entries", {})
            except (OSError, ValueError):
                self.entries = {}

    def save(self):
        os.makedirs(self.directory, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=self.directory, prefix=".tuning")
        with os.fdopen(fd, "w") as f:
            json.dump({"version": CANDIDATE_SET_VERSION, "entries": self.entries}, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def lookup(self, key, buckets, precision):
        """(tuning, missing decisions): the cached decisions for this plan's buckets."""
        t = _lib.Tuning()
        entry = self.entries.get(key, {})
        missing = []
        for kind in KINDS:
            if kind in ("p2p", "stage") and precision != "mixed":
                continue
            table = entry.get(kind, {})
            for level, b in buckets[kind].items():
                v = table.get(str(b))
                if v is None:
                    missing.append((kind, level))
                    continue
                if kind == "p2p":
                    t.p2p_uniform_min = int(v)
                elif kind == "stage":
                    t.m2l_stage = int(v)
                elif kind == "m2l":
                    t.m2l_mode[level] = int(v)
                elif kind == "l2l":
                    t.l2l_width[level] = int(v)
                else:
                    t.m2m_form[level] = int(v)
        return t, missing

    def store(self, key, buckets, tuning, report, precision):
        """Keep the decisions whose static candidate was timed (report not NaN)."""
        entry = self.entries.setdefault(key, {})
        for k, kind in enumerate(KINDS):
            if kind in ("p2p", "stage") and precision != "mixed":
                continue
            table = entry.setdefault(kind, {})
            for level, b in buckets[kind].items():
                lv = 0 if kind in ("p2p", "stage") else level
                if level >= 17 or math.isnan(report[k][lv][0]):
                    continue
                if kind == "p2p":
                    v = tuning.p2p_uniform_min
                elif kind == "stage":
                    v = tuning.m2l_stage
                elif kind == "m2l":
                    v = tuning.m2l_mode[level]
                elif kind == "l2l":
                    v = tuning.l2l_width[level]
                else:
                    v = tuning.m2m_form[level]
                table[str(b)] = int(v)


def autotune(plan, p, precision, gradients=True, device=0, m2l="dense", budget=10.0):
    """fmmh_autotune on this plan: (tuning, report[kind][level][candidate] in ms; kinds
    P2P, M2L, L2L, M2M, M2L stage)."""
    from . import M2L_METHODS, PRECISIONS

    o = _lib.OperatorOptions(int(p), PRECISIONS[precision], int(bool(gradients)), int(device), M2L_METHODS[m2l])
    t = _lib.Tuning()
    report = np.full(5 * 17 * 5, np.nan)
    _lib.check(_lib.lib().fmmh_autotune(plan.handle, C.byref(o), float(budget), C.byref(t),
                                        C.c_void_p(report.ctypes.data)))
    return t, report.reshape(5, 17, 5)


def resolve(plan, p, precision, gradients, device, m2l, directory, budget):
    """The tuning of an operator over `plan`: cached decisions for its buckets, the missing
    ones tuned within `budget` seconds and stored.  Returns (tuning, info)."""
    cache = TuningCache(directory)
    key = cache_key(p, precision, gradients, m2l, device)
    buckets = level_buckets(plan)
    t, missing = cache.lookup(key, buckets, precision)
    info = {"key": key, "cached": not missing, "tuned_s": 0.0}
    if missing and budget > 0:
        t0 = time.time()
        tuned, report = autotune(plan, p, precision, gradients, device, m2l, budget)
        info["tuned_s"] = time.time() - t0
        info["report_ms"] = report
        cache.store(key, buckets, tuned, report, precision)
        cache.save()
        # the decisions as the cache holds them (levels of one bucket share one decision,
        # the last level timed), so this build and later ones use the same shapes
        t, _ = cache.lookup(key, buckets, precision)
    info["decisions"] = describe(t, plan.nlevels)
    return t, info


def describe(t, nlevels) -> dict:
    return {"p2p_uniform_min": int(t.p2p_uniform_min), "m2l_stage": int(t.m2l_stage),
            "m2l_mode": [int(t.m2l_mode[l]) for l in range(min(nlevels, 17))],
            "l2l_width": [int(t.l2l_width[l]) for l in range(min(nlevels, 17))],
            "m2m_form": [int(t.m2m_form[l]) for l in range(min(nlevels, 17))]}


def tune_leaf_size(builder, points, candidates, directory=None, repeats=3):
    """The max points per leaf among `candidates` whose FMM evaluates `points` fastest
    (unit charges, whole evaluations on the device after one warm-up), cached by
    (key, N bucket) when a directory is given.  Returns (leaf, {leaf: ms})."""
    from . import Operator

    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    key = cache_key(builder._p, builder._precision, builder._gradients, builder._m2l, builder._device)
    nb = str(bucket(len(pts)))
    cache = TuningCache(directory) if directory is not None else None
    if cache is not None:
        v = cache.entries.get(key, {}).get("leaf", {}).get(nb)
        if v is not None and int(v) in candidates:
            return int(v), {}
    times = {}
    for leaf in candidates:
        plan = builder._with_leaf(int(leaf)).build_plan(pts)
        op = Operator(plan, builder._p, builder._precision, builder._gradients, builder._device, builder._m2l)
        op.load_charges(np.ones(len(pts)))
        op.reset()
        op.evaluate()
        op.sync()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            op.reset()
            op.evaluate()
            op.sync()
            ts.append(time.perf_counter() - t0)
        times[int(leaf)] = float(np.median(ts)) * 1e3
        del op, plan
    best = min(times, key=lambda k: (times[k], list(candidates).index(k)))
    if cache is not None:
        cache.entries.setdefault(key, {}).setdefault("leaf", {})[nb] = best
        cache.save()
    return best, times
