"""Benchmark of the FMM evaluation hot path on one MI355X.

Workload: the reference's GPU configuration -- N = 4,000,000 points uniform in the unit
cube from SplitMix64(0x00be9c06) (fmm-bench/src/measure.rs:35-58), charges N(0,1) from the
same stream (Box-Muller), sources = targets, potential and gradient, at most 128 points
per leaf, degree p (12 by default: relative error of the potential below 1e-6).

One step = one evaluation with the charges already in HBM: charge gather, reset, every
operator in the evaluator's order, and the output pass into device buffers.  The
mixed-precision path is the headline value; the plain f64 path is timed beside it in
the same run, and the error of both against the f64 direct sum at 4,096 sampled targets
is reported.  The oracle's serial CPU FMM on a bounded sample is the CPU baseline.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_F64_MFMA = 78.6  # TFLOP/s, dense FP64 matrix
PEAK_F32 = 157.3  # TFLOP/s, FP32 (vector and dense matrix)
PEAK_BF16 = 2500.0  # TFLOP/s, dense BF16 MFMA
PEAK_HBM = 8000.0  # GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_laplace4M_p12_mixed.txt")


def hbm_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the default workload (scaled by the copy-probe factors)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    scale, rows = {}, {}
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] in ("fetch_scale", "write_scale"):
            scale[f[0]] = float(f[1])
        else:
            rows[f[0]] = (float(f[1]), float(f[2]))
    if kernel not in rows:
        return None
    fk, wk = rows[kernel]
    return fk * 1024 * scale.get("fetch_scale", 1.0) + wk * 1024 * scale.get("write_scale", 1.0)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, pts, q, sample, near, exact_phi, exact_grad):
    """The reference's CPU path restated (oracle/cpu_path.c: OpenMP threads over targets,
    AVX-512 targets-in-lanes P2P with vrsqrt14 + two Newton steps, the batched M2L column
    kernel) on this host's cores, in this run: configuration 1 (100,000 uniform points,
    p = 6, <= 64 per leaf, charges uniform in [-1, 1)) and one evaluation of the
    benchmark workload itself (same points, charges, degree, leaf size and near rule).
    Tree and table builds are excluded from the times, as for the GPU."""
    from oracle import oracle as O  # the CPU baseline only

    threads = O.cpu_threads()
    # configuration 1
    p1, q1, s1 = O.bench_problem(100_000, 1000)
    f1 = O.Fmm(p1, p=6, leaf_points=64)
    t1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        phi1, grad1 = f1.evaluate_cpu_path(q1)
        t1.append(time.perf_counter() - t0)
    e1p, e1g = O.direct_sum(p1, q1, p1[s1])
    c1 = {"n": 100_000, "p": 6, "leaf_points": 64, "ms": min(t1) * 1e3,
          "mpts_per_s": 100_000 / min(t1) / 1e6,
          "err_phi": O.error_norms(phi1[s1], e1p / (4 * np.pi))[0],
          "err_grad": O.error_norms(grad1[s1], e1g / (4 * np.pi), True)[0]}
    del f1
    # the benchmark workload
    t0 = time.perf_counter()
    f = O.Fmm(pts, p=args.p, leaf_points=args.leaf, near_rule=near)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    phi, grad = f.evaluate_cpu_path(q)
    t_eval = time.perf_counter() - t0
    err_phi = float(np.linalg.norm(phi[sample] - exact_phi) / np.linalg.norm(exact_phi))
    err_grad = float(np.linalg.norm(grad[sample] - exact_grad) / np.linalg.norm(exact_grad))
    del f
    n = len(pts)
    return {"value": n / t_eval / 1e6, "unit": "Mpts/s", "cores": threads, "kind": "port",
            "sample": f"restatement of the reference CPU path (oracle/cpu_path.c, {O.cpu_isa()}, {threads} OpenMP "
                      f"threads) on the whole benchmark workload: N={n}, p={args.p}, <= {args.leaf} per leaf, one "
                      f"evaluation {t_eval:.2f} s (tree and tables {t_build:.1f} s excluded)",
            "ms_per_eval": t_eval * 1e3, "err_phi": err_phi, "err_grad": err_grad,
            "cpu_model": cpu_model(), "isa": O.cpu_isa(), "nproc": os.cpu_count(),
            "config1": c1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--p", type=int, default=12)
    ap.add_argument("--leaf", type=int, default=128)
    ap.add_argument("--sampled", type=int, default=4096)
    ap.add_argument("--near", default="device", help="near rule: off, device, or N_W,N_X")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain f64 leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 256 MiB device to device (a known-size kernel for counter scaling)")
    ap.add_argument("--m2l", default="dense", choices=("dense", "rotation", "plane_wave"),
                    help="M2L formulation of both legs (rotation: f32 in the mixed leg; plane_wave: f64 waves by "
                         "merge and shift in both legs)")
    ap.add_argument("--workload", default="cube", choices=("cube", "batched", "helmholtz", "nonuniform", "grid"),
                    help="cube: 4M uniform points (the headline); batched: 256 independent Gaussian clouds, "
                         "p = 6, one forest evaluation replayed as a hipGraph, plus its adjoint; helmholtz: "
                         "1M uniform points, k = 25, complex charges, degree by the rule at 1e-6; nonuniform: 2M "
                         "sphere-surface + 1M Plummer points, p = 10, rotation M2L in f32; grid: the reference's published "
                         "device grid (N = 1e5..1e7, p = 3..12, mixed and f64) with its published times and errors")
    ap.add_argument("--k", type=float, default=25.0, help="helmholtz: the wavenumber (real)")
    ap.add_argument("--tune-budget", type=float, default=10.0,
                    help="seconds of autotune per leg (fmm_amd/tune.py) before timing; 0: the static rule")
    ap.add_argument("--tuning-cache", default=None,
                    help="directory of the tuning cache (default: a fresh temporary directory, so each run tunes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return batched(args)
    if args.workload == "helmholtz":
        return helmholtz(args)
    if args.workload == "grid":
        return grid(args)
    if args.workload == "nonuniform":
        return nonuniform(args)

    import torch

    import fmm_amd
    from fmm_amd import points as P

    dev = torch.device("cuda:0")
    if args.copy_probe:
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 268435456 bytes read, 268435456 bytes written")
        del a, b
    t0 = time.time()
    pts, q, sample = P.bench_cube(args.n, args.sampled)
    log(f"problem: N={args.n} drawn in {time.time() - t0:.1f}s")

    if args.near == "off":
        near = None
    elif args.near == "device":
        near = (0.5 * (args.p + 1) ** 2,) * 2
    else:
        near = tuple(float(v) for v in args.near.split(","))
    t0 = time.time()
    plan = fmm_amd.Plan(pts, leaf_points=args.leaf, near_rule=near)
    t_plan = time.time() - t0
    lk = plan.leaf_keys()
    hist = {int(l): int(c) for l, c in zip(*np.unique(lk & np.uint64(0x7FFF), return_counts=True))}
    log(f"plan: {plan.nlevels} levels, {plan.nleaves} leaves {hist}, built in {t_plan:.1f}s")

    q_dev = torch.from_numpy(q).to(dev)
    phi_dev = torch.zeros(args.n, dtype=torch.float64, device=dev)
    grad_dev = torch.zeros(3 * args.n, dtype=torch.float64, device=dev)

    t0 = time.time()
    exact_phi, exact_grad = fmm_amd.direct_sum(pts, q, pts[sample])  # compensated f64, on the GPU
    exact_phi /= 4 * np.pi
    exact_grad /= 4 * np.pi
    log(f"direct sum at {len(sample)} targets in {time.time() - t0:.1f}s")

    def error_norm(a, b):
        """Relative L2 error (fmm-validate/src/metrics.rs:51-170)."""
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    import tempfile

    from fmm_amd import tune

    tune_dir = args.tuning_cache or tempfile.mkdtemp(prefix="fmm_amd_tune")

    def run(precision):
        m2l = args.m2l
        tuning, tinfo = None, None
        if args.tune_budget > 0:
            tuning, tinfo = tune.resolve(plan, args.p, precision, True, 0, m2l, tune_dir, args.tune_budget)
            log(f"[{precision}] tuning ({'cached' if tinfo['cached'] else 'tuned in %.1fs' % tinfo['tuned_s']}): "
                f"{tinfo['decisions']}")
        t0 = time.time()
        op = fmm_amd.Operator(plan, args.p, precision, True, 0, m2l, tuning)
        stats = op.stats()
        log(f"[{precision}] operator built in {time.time() - t0:.1f}s: {stats}")

        def step():
            op.load_charges(q_dev)
            op.reset()
            op.evaluate()
            op.read_output_device(phi_dev, grad_dev)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            s = time.perf_counter()
            step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - s)
        # the errors of the timed (eager) steps' own output
        phi = phi_dev.cpu().numpy()
        grad = grad_dev.cpu().numpy().reshape(-1, 3)
        e_phi = error_norm(phi[sample], exact_phi)
        e_grad = error_norm(grad[sample], exact_grad)
        # per-kind device times in a pass of their own (HIP events on the operator's own
        # stream around each launch; kept out of the timed steps above)
        op.kind_timing(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        kinds = op.kind_times()
        op.kind_timing(False)
        # the same steps with reset + evaluate replayed as one hipGraph launch
        def step_graph():
            op.load_charges(q_dev)
            op.evaluate_graph()
            op.read_output_device(phi_dev, grad_dev)

        step_graph()
        torch.cuda.synchronize()
        gtimes = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            s = time.perf_counter()
            step_graph()
            torch.cuda.synchronize()
            gtimes.append(time.perf_counter() - s)
        # the graph replay's output against the eager output (same launches, same bits)
        graph_diff = float(np.abs(phi_dev.cpu().numpy() - phi).max())
        med = float(np.median(times))
        log(f"[{precision}] median {med * 1e3:.2f} ms ({args.n / med / 1e6:.1f} Mpts/s), "
            f"err phi {e_phi:.3e} grad {e_grad:.3e}")
        log(f"[{precision}] kinds (ms, launches): " +
            ", ".join(f"{k} {v[0] / max(v[1], 1):.3f}x{v[1]}" for k, v in kinds.items() if v[1]))
        gmed = float(np.median(gtimes))
        log(f"[{precision}] hipGraph replay: median {gmed * 1e3:.2f} ms ({args.n / gmed / 1e6:.1f} Mpts/s), "
            f"max |graph - eager| phi {graph_diff:.1e}")
        res = dict(ms=med * 1e3, times=times, mpts=args.n / med / 1e6, err_phi=e_phi, err_grad=e_grad,
                   kinds=kinds, stats=stats, graph_ms=gmed * 1e3, graph_diff=graph_diff,
                   tuning=None if tinfo is None else {"decisions": tinfo["decisions"],
                                                      "tuned_s": round(tinfo["tuned_s"], 2),
                                                      "cached": tinfo["cached"]})
        del op
        return res

    mixed = run("mixed")
    plain = None if args.no_plain else run("f64")

    # roofline of the dominant kernel of the headline (mixed) path; `traffic` (HBM bytes
    # per launch) comes from the committed counter runs of the default workload
    default_workload = (args.n, args.p, args.leaf, args.near) == (4_000_000, 12, 128, "device")
    stats = mixed["stats"]
    kinds = mixed["kinds"]
    n = (args.p + 1) ** 2
    steps = max(1, args.steps)
    per_step = {k: v[0] / steps for k, v in kinds.items() if v[1]}
    contract_ms = per_step.get("m2l_contract", 0.0)
    p2p_ms = per_step.get("p2p", 0.0)
    if contract_ms >= p2p_ms:
        # one launch per evaluation covers every level's M2L
        flops = 2.0 * n * n * stats["m2l_pairs"]
        avg_ms = kinds["m2l_contract"][0] / kinds["m2l_contract"][1]
        achieved = flops / (avg_ms * 1e-3) / 1e12
        roof = {"kernel": "k_contract_bf16 (M2L, bf16x3 split)", "bound": "mfma", "achieved": achieved,
                "peak": PEAK_BF16, "unit": "TFLOP/s", "frac": achieved / PEAK_BF16,
                "traffic": hbm_traffic("k_contract_bf16") if default_workload else None,
                "algorithmic": "2 (p+1)^4 flops per V-list pair x m2l_pairs per launch; the bf16x3 split "
                               "issues 6 bf16 products per algorithmic product",
                "executed_bf16_tflops": 6 * achieved, "executed_frac": 6 * achieved / PEAK_BF16,
                "flops_per_launch": flops}
    else:
        pairs = stats["p2p_pairs"]
        avg_ms = kinds["p2p"][0] / kinds["p2p"][1]
        achieved = 30.0 * pairs / (avg_ms * 1e-3) / 1e12
        roof = {"kernel": "k_p2p<float> (P2P)", "bound": "valu", "achieved": achieved, "peak": PEAK_F32,
                "unit": "TFLOP/s", "frac": achieved / PEAK_F32,
                "traffic": hbm_traffic("k_p2p") if default_workload else None,
                "algorithmic": "15 lane-ops (30 flops) per source-target pair x p2p_pairs per launch, FP32 peak",
                "compute_unit": "FP32 vector ALU (no matrix cores: the bound is arithmetic, not HBM)",
                "flops_per_launch": 30.0 * pairs}
    # the other large kernel beside it
    c_ms = kinds["m2l_contract"][0] / max(1, kinds["m2l_contract"][1]) if "m2l_contract" in kinds else 0.0
    if c_ms > 0:
        c_flops = 2.0 * n * n * stats["m2l_pairs"]
        c_tf = c_flops / (c_ms * 1e-3) / 1e12
        roof["m2l_contract"] = {"avg_launch_ms": c_ms, "achieved": c_tf, "executed_bf16_tflops": 6 * c_tf,
                                "executed_frac_bf16": 6 * c_tf / PEAK_BF16,
                                "traffic": hbm_traffic("k_contract_bf16") if default_workload else None}
    roof["avg_launch_ms"] = avg_ms

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, pts, q, sample, near, exact_phi, exact_grad)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "Laplace FMM Mpts/s at 1e-6 rel err: mixed-prec batched vs plain FP64 GPU",
        "value": mixed["mpts"],
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 P2P (fixed-point separations) + bf16x3 M2L (f32 MFMA accumulation) + f64 other",
        "data": "synthetic: SplitMix64(0x00be9c06) uniform cube points, N(0,1) charges",
        "config": {"workload": "laplace-fmm-uniform-cube", "n": args.n, "p": args.p, "leaf_points": args.leaf,
                   "m2l": args.m2l,
                   "charges": "normal", "gradients": True, "near_rule": near, "sampled_targets": len(sample),
                   "leaves": int(plan.nleaves), "levels": int(plan.nlevels), "leaf_levels": hist},
        "errors": {"mixed": {"phi": mixed["err_phi"], "grad": mixed["err_grad"]}},
        "graph_ms_per_step": mixed["graph_ms"],
        "kind_ms": {k: round(v, 4) for k, v in per_step.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if plain is not None:
        line["plain_f64"] = {"value": plain["mpts"], "ms_per_step": plain["ms"], "err_phi": plain["err_phi"],
                             "err_grad": plain["err_grad"],
                             "kind_ms": {k: round(v[0] / steps, 4) for k, v in plain["kinds"].items() if v[1]}}
        line["speedup_vs_plain_f64"] = mixed["mpts"] / plain["mpts"]
        line["errors"]["plain_f64"] = {"phi": plain["err_phi"], "grad": plain["err_grad"]}
    line["p2p_pairs"] = stats["p2p_pairs"]
    line["m2l_pairs"] = stats["m2l_pairs"]
    # interactions per second of the timed evaluation (BASELINE's definitions)
    line["interactions"] = {"p2p_pairs_per_s": stats["p2p_pairs"] / (mixed["ms"] * 1e-3),
                            "m2l_v_pairs_per_s": stats["m2l_pairs"] / (mixed["ms"] * 1e-3)}
    line["graph_max_abs_diff_phi"] = mixed["graph_diff"]
    line["tuning"] = {"mixed": mixed["tuning"], "plain_f64": None if plain is None else plain["tuning"],
                      "budget_s": args.tune_budget}
    print(json.dumps(line), flush=True)


def batched(args):
    """256 independent clouds (sizes log-uniform in [2k, 256k], one Gaussian blob each,
    charges N(0,1)), p = 6, f32 P2P + bf16x3 M2L, <= 64 points per leaf, near rule at
    (p+1)^2 / 2: one forest
    operator, reset + evaluate replayed as one hipGraph launch; then the adjoint of
    L = sum(phi^2)/N (one more evaluation and the combine kernel)."""
    import torch

    import fmm_amd
    from fmm_amd import points as P

    dev = torch.device("cuda:0")
    t0 = time.time()
    clouds, charges = P.batched_clouds()
    n = int(sum(len(c) for c in clouds))
    log(f"batched: {len(clouds)} clouds, {n} points, drawn in {time.time() - t0:.1f}s")
    t0 = time.time()
    # the near rule on, as the reference's FmmBuilder has it by default (fmm-exec/src/fmm.rs:1029),
    # at this path's crossovers (p+1)^2 / 2
    builder = fmm_amd.FmmBuilder(6).precision("mixed").max_points_per_leaf(64).near_rule("device")
    if args.tune_budget > 0:
        import tempfile

        builder.tuning_cache(args.tuning_cache or tempfile.mkdtemp(prefix="fmm_amd_tune")).tuning_budget(args.tune_budget)
    fmm = builder.build_batched(clouds)
    if builder.tuning_info is not None:
        log(f"tuning: {builder.tuning_info['decisions']} ({builder.tuning_info['tuned_s']:.1f} s)")
    log(f"forest plan + operator in {time.time() - t0:.1f}s: {fmm.operator.stats()}")
    op = fmm.operator
    q_dev = torch.from_numpy(np.concatenate(charges)).to(dev)
    phi = torch.zeros(n, dtype=torch.float64, device=dev)
    field = torch.zeros(n, 3, dtype=torch.float64, device=dev)

    def step():
        op.load_charges(q_dev)
        op.evaluate_graph()
        op.read_output_device(phi, field)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        s = time.perf_counter()
        step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - s)
    med = float(np.median(times))
    # adjoint of L = sum(phi^2) / N on the device
    g = (2.0 / n) * phi
    gq = torch.empty_like(phi)
    gz = torch.empty_like(field)
    vt = []
    for _ in range(max(2, min(args.steps, 5))):
        torch.cuda.synchronize()
        s = time.perf_counter()
        op.vjp_device(q_dev, g, field, gq, gz)
        torch.cuda.synchronize()
        vt.append(time.perf_counter() - s)
    gq_h, gz_h = gq.cpu().numpy(), gz.cpu().numpy()
    step()
    # the adjoint against central finite differences (h = 1e-4) of the f64 direct-sum loss
    # L = sum(phi^2) / N over the forest; clouds do not interact, so a coordinate of the
    # smallest cloud moves only that cloud's terms (direct sums on the GPU, compensated)
    first = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    cs = int(np.argmin([len(c) for c in clouds]))
    zc, qc = clouds[cs], charges[cs]
    gqc, gzc = gq_h[first[cs]:first[cs + 1]], gz_h[first[cs]:first[cs + 1]]

    def loss(zz, qq):
        f, _ = fmm_amd.direct_sum(zz, qq, zz, gradients=False)
        return float(np.sum((f / (4 * np.pi)) ** 2) / n)

    hstep, fd_worst = 1e-4, 0.0
    for j in P.SplitMix64(77).sample(len(zc), 4):
        qp, qm = qc.copy(), qc.copy()
        qp[j] += hstep
        qm[j] -= hstep
        fd = (loss(zc, qp) - loss(zc, qm)) / (2 * hstep)
        fd_worst = max(fd_worst, abs(fd - gqc[j]) / np.abs(gqc).max())
        for k in range(3):
            zp, zm = zc.copy(), zc.copy()
            zp[j, k] += hstep
            zm[j, k] -= hstep
            fd = (loss(zp, qc) - loss(zm, qc)) / (2 * hstep)
            fd_worst = max(fd_worst, abs(fd - gzc[j, k]) / np.abs(gzc).max())
    log(f"adjoint vs finite differences (cloud {cs}, {len(zc)} points): worst {fd_worst:.2e}")
    # errors: 32 sampled targets per cloud against the compensated f64 direct sum
    ph = phi.cpu().numpy()
    gr = field.cpu().numpy()
    num = den = numg = deng = 0.0
    for c, (pts, q) in enumerate(zip(clouds, charges)):
        idx = P.SplitMix64(c + 1).sample(len(pts), 32)
        e, eg = fmm_amd.direct_sum(pts, q, pts[idx])
        e /= 4 * np.pi
        eg /= 4 * np.pi
        num += float(np.sum((ph[first[c] + idx] - e) ** 2))
        den += float(np.sum(e ** 2))
        numg += float(np.sum((gr[first[c] + idx] - eg) ** 2))
        deng += float(np.sum(eg ** 2))
    line = {
        "metric": "batched Laplace FMM Mpts/s (256 clouds, one forest, hipGraph replay)",
        "value": n / med / 1e6, "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": med * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 P2P (fixed-point separations) + bf16x3 M2L + f64 other", "data": "synthetic",
        "config": {"workload": "batched-gaussian-clouds", "clouds": len(clouds), "n": n, "p": 6, "leaf_points": 64,
                   "near_rule": [24.5, 24.5], "sizes": [int(len(c)) for c in clouds][:8] + ["..."]},
        "errors": {"phi": (num / den) ** 0.5, "grad": (numg / deng) ** 0.5, "sampled_per_cloud": 32},
        "adjoint_ms": float(np.median(vt)) * 1e3,
        "adjoint_fd": {"worst_rel_to_max": fd_worst, "h": hstep, "cloud_points": int(len(zc)),
                       "checked": "dL/dq and dL/dz at 4 points of the smallest cloud, central differences of the "
                                  "f64 direct-sum loss sum(phi^2)/N"},
        "tuning": None if builder.tuning_info is None else builder.tuning_info["decisions"],
    }
    print(json.dumps(line), flush=True)


def nonuniform(args):
    """The non-uniform configuration: 2,000,000 points on the sphere surface + 1,000,000
    Plummer points (a = 0.1, truncated at 10a) in the unit cube, charges uniform in [0, 1),
    p = 10, deep adaptive octree (max level 16), at most 128 points per leaf, near rule at
    the device crossovers; rotation M2L in f32 with f64 M2M/L2L and leaf operators (the
    configuration's method), beside the dense bf16x3 and the plain f64 paths.  Reports the
    leaf-level histogram and the W/X pair counts; errors against the f64 direct sum at 4,096
    sampled targets."""
    import torch

    import fmm_amd
    from fmm_amd import points as P

    p, leaf = (args.p if args.p != 12 else 10), args.leaf
    t0 = time.time()
    pts, q = P.nonuniform()
    n = len(pts)
    sample = P.SplitMix64(P.BENCH_SEED + 3).sample(n, args.sampled)
    log(f"nonuniform: N={n} drawn in {time.time() - t0:.1f}s")
    near = (0.5 * (p + 1) ** 2,) * 2
    t0 = time.time()
    plan = fmm_amd.Plan(pts, leaf_points=leaf, near_rule=near)
    lk = plan.leaf_keys()
    hist = {int(l): int(c) for l, c in zip(*np.unique(lk & np.uint64(0x7FFF), return_counts=True))}
    wx = {"w": int(sum(len(plan.list("w", l)[1]) for l in range(plan.nlevels))),
          "x": int(sum(len(plan.list("x", l)[1]) for l in range(plan.nlevels)))}
    log(f"plan: {plan.nlevels} levels, leaves {hist}, W/X pairs {wx}, built in {time.time() - t0:.1f}s")
    ep, eg = fmm_amd.direct_sum(pts, q, pts[sample])
    ep, eg = ep / (4 * np.pi), eg / (4 * np.pi)
    dev = torch.device("cuda:0")
    q_dev = torch.from_numpy(q).to(dev)
    phi = torch.zeros(n, dtype=torch.float64, device=dev)
    grad = torch.zeros(3 * n, dtype=torch.float64, device=dev)

    def run(precision, m2l):
        op = fmm_amd.Operator(plan, p, precision, True, 0, m2l)
        log(f"[{precision}, {m2l}] operator: {op.stats()}")

        def step():
            op.load_charges(q_dev)
            op.reset()
            op.evaluate()
            op.read_output_device(phi, grad)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            s = time.perf_counter()
            step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - s)
        # the errors of the timed (eager) steps' own output
        ph = phi.cpu().numpy()
        gr = grad.cpu().numpy().reshape(-1, 3)
        err = {"phi": float(np.linalg.norm(ph[sample] - ep) / np.linalg.norm(ep)),
               "grad": float(np.linalg.norm(gr[sample] - eg) / np.linalg.norm(eg))}
        # per-kind device times in a pass of their own, outside the timed steps
        op.kind_timing(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        kinds = op.kind_times()
        op.kind_timing(False)
        med = float(np.median(times))
        per = {k: round(v[0] / args.steps, 3) for k, v in kinds.items() if v[1]}
        log(f"[{precision}, {m2l}] {med * 1e3:.2f} ms ({n / med / 1e6:.1f} Mpts/s), errors {err}, kinds {per}")
        return {"ms_per_step": med * 1e3, "value": n / med / 1e6, "errors": err, "kind_ms": per}

    rot = run("mixed", "rotation")
    dense = run("mixed", "dense")
    plain = run("f64", "dense")
    pw = run("mixed", "plane_wave")
    line = {
        "metric": "non-uniform Laplace FMM Mpts/s (2M sphere + 1M Plummer, p = 10, rotation M2L f32)",
        "value": rot["value"], "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rot["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 P2P + f32 rotation M2L + f64 M2M/L2L and leaf operators", "data": "synthetic",
        "config": {"workload": "nonuniform-sphere-plummer", "n": n, "p": p, "leaf_points": leaf, "max_level": 16,
                   "near_rule": near, "leaf_levels": hist, "wx_pairs": wx, "levels": int(plan.nlevels)},
        "errors": rot["errors"], "kind_ms": rot["kind_ms"],
        "mixed_dense_bf16x3": dense, "plain_f64_dense": plain, "mixed_plane_wave_f64": pw,
    }
    print(json.dumps(line), flush=True)


def helmholtz(args):
    """The Helmholtz configuration: N = 1,000,000 uniform in the unit cube, k = 25 (about
    four wavelengths across), complex charges with N(0,1) parts, degree by the rule at
    1e-6 (fmm-exec/src/helmholtz/degree.rs), at most 256 points per leaf (the reference's
    default at 1e-6), potential and gradient.  Plain f64 and mixed (complex64 P2P with
    complex128 sums, bf16x3 M2L on MFMA) both timed; errors against the compensated f64
    direct sum (on the GPU) at 1,000 sampled targets; the oracle's serial Helmholtz FMM on
    a bounded sample is the CPU baseline."""
    import torch

    import fmm_amd
    from fmm_amd import points as P

    n = args.n if args.n != 4_000_000 else 1_000_000
    leaf = args.leaf if args.leaf != 128 else 256
    dev = torch.device("cuda:0")
    t0 = time.time()
    pts, q, sample = P.helmholtz_cube(n, 1000)
    log(f"helmholtz: N={n} k={args.k} drawn in {time.time() - t0:.1f}s")
    t0 = time.time()
    b = fmm_amd.HelmholtzFmmBuilder(args.k, tolerance=1e-6)
    p = b.degree_for(fmm_amd.Plan(pts, leaf_points=leaf))
    # W and X pairs to P2P at this path's crossovers, w = x = (p+1)^2 (WxNear::Crossovers)
    near = None if args.near == "off" else (float((p + 1) ** 2),) * 2
    plan = fmm_amd.Plan(pts, leaf_points=leaf, near_rule=near)
    lk = plan.leaf_keys()
    hist = {int(l): int(c) for l, c in zip(*np.unique(lk & np.uint64(0x7FFF), return_counts=True))}
    log(f"plan: {plan.nlevels} levels, leaves {hist}, p = {p}, built in {time.time() - t0:.1f}s")
    t0 = time.time()
    ep, eg = fmm_amd.helmholtz_direct_sum(args.k, pts, q, pts[sample])
    ep, eg = ep / (4 * np.pi), eg / (4 * np.pi)
    log(f"direct sum at {len(sample)} targets in {time.time() - t0:.1f}s")
    q_dev = torch.from_numpy(np.ascontiguousarray(q, dtype=np.complex128)).to(dev)

    def run(precision, m2l="dense"):
        t0 = time.time()
        op = fmm_amd.HelmholtzOperator(plan, p, args.k, precision, True, 0, m2l)
        log(f"{precision}, {m2l} M2L: operator built in {time.time() - t0:.1f}s: {op.stats()}")

        def step():
            op.load_charges(q_dev)
            op.reset()
            op.evaluate()

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            s = time.perf_counter()
            step()
            torch.cuda.synchronize()  # the device, so the operator's own stream too
            times.append(time.perf_counter() - s)
        op.kind_timing(True)
        step()
        kinds = op.kind_times()
        op.kind_timing(False)
        out = op.read_output()
        err = {"phi": float(np.linalg.norm(out.potential[sample] - ep) / np.linalg.norm(ep)),
               "grad": float(np.linalg.norm(out.gradient[sample] - eg) / np.linalg.norm(eg))}
        med = float(np.median(times))
        log(f"{precision}, {m2l} M2L: {med * 1e3:.2f} ms/step, errors {err}, kinds "
            + ", ".join(f"{k} {v[0]:.2f}" for k, v in kinds.items() if v[1]))
        return med, err, kinds, op.stats()

    mixed = run("mixed")
    plain = run("f64")
    rot = run("f64", "rotation")
    pw = run("f64", "plane_wave")
    pw_mixed = run("mixed", "plane_wave")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        m = 6000
        t0 = time.time()
        f = O.Fmm(pts[:m], p=p, leaf_points=leaf)
        f.evaluate_helmholtz(args.k, q[:m])
        dt = time.time() - t0
        cpu = {"value": m / dt / 1e6, "unit": "Mpts/s", "cores": 1, "kind": "port",
               "sample": f"oracle serial Helmholtz FMM (tree, tables and evaluation) on the first {m} points, "
                         f"p = {p}, leaf {leaf}"}
    # the headline leg: the configuration's M2L (the exponential M2L by merge and shift, as
    # the reference runs configuration 4) with the mixed P2P
    med, err, kinds, stats = pw_mixed
    p2p_ms = kinds["p2p"][0]
    line = {
        "metric": "Helmholtz FMM Mpts/s (1M uniform, k = 25, 1e-6 degree rule, complex charges)",
        "value": n / med / 1e6, "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": med * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64 P2P (complex128 sums) + f64 exponential M2L + f64 other", "data": "synthetic",
        "config": {"workload": "helmholtz-uniform-cube", "n": n, "k": args.k, "p": p, "leaf_points": leaf,
                   "near_wx": near, "m2l": "plane_wave"},
        "errors": err,
        "mixed_dense_m2l": {"ms_per_step": mixed[0] * 1e3, "value": n / mixed[0] / 1e6, "errors": mixed[1],
                            "m2l_ms": mixed[2]["m2l"][0], "device_bytes": mixed[3]["device_bytes"]},
        "plain_f64": {"ms_per_step": plain[0] * 1e3, "value": n / plain[0] / 1e6, "errors": plain[1],
                      "device_bytes": plain[3]["device_bytes"]},
        "plain_f64_rotation_m2l": {"ms_per_step": rot[0] * 1e3, "value": n / rot[0] / 1e6, "errors": rot[1],
                                   "m2l_ms": rot[2]["m2l"][0], "device_bytes": rot[3]["device_bytes"]},
        # the exponential M2L (merge and shift, f64) on every level the reference's rule
        # gives a quadrature: the configuration's M2L in the reference
        "plain_f64_plane_wave_m2l": {"ms_per_step": pw[0] * 1e3, "value": n / pw[0] / 1e6, "errors": pw[1],
                                     "m2l_ms": pw[2]["m2l"][0], "device_bytes": pw[3]["device_bytes"]},

        "device_bytes": stats["device_bytes"],
        "kinds_ms": {k: v[0] for k, v in kinds.items() if v[1]},
        "interactions_per_s": stats["p2p_pairs"] / (p2p_ms * 1e-3) if p2p_ms > 0 else None,
        "m2l_pairs": stats["m2l_pairs"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# The reference's published device grid on the GH200's H100 (fmm-bench/results/
# phase4s-gh200.md:89-113), per (precision, N, p): median ms per evaluation, error phi,
# error grad phi (relative L2 at 1,000 sampled targets against the f64 direct sum).
PUBLISHED_GRID = {
    ("f32", 100000, 3): (1.572, 3.39e-3, 1.99e-3), ("f32", 100000, 6): (1.889, 1.42e-4, 1.30e-4),
    ("f32", 100000, 8): (2.657, 2.12e-5, 2.37e-5), ("f32", 1000000, 3): (12.755, 2.00e-3, 3.80e-3),
    ("f32", 1000000, 6): (15.917, 8.53e-5, 2.56e-4), ("f32", 1000000, 8): (21.599, 1.36e-5, 4.59e-5),
    ("f32", 10000000, 3): (220.777, 1.87e-3, 1.87e-3), ("f32", 10000000, 6): (257.726, 7.90e-5, 1.23e-4),
    ("f32", 10000000, 8): (315.430, 1.21e-5, 2.26e-5),
    ("f64", 100000, 3): (1.940, 3.39e-3, 1.99e-3), ("f64", 100000, 6): (2.555, 1.42e-4, 1.30e-4),
    ("f64", 100000, 8): (3.785, 2.12e-5, 2.37e-5), ("f64", 100000, 12): (8.492, 8.07e-7, 1.11e-6),
    ("f64", 1000000, 3): (20.560, 2.00e-3, 3.80e-3), ("f64", 1000000, 6): (25.475, 8.53e-5, 2.56e-4),
    ("f64", 1000000, 8): (35.268, 1.36e-5, 4.59e-5), ("f64", 1000000, 12): (76.389, 4.35e-7, 2.17e-6),
    ("f64", 10000000, 3): (361.077, 1.87e-3, 1.87e-3), ("f64", 10000000, 6): (415.426, 7.90e-5, 1.23e-4),
    ("f64", 10000000, 8): (517.091, 1.21e-5, 2.26e-5), ("f64", 10000000, 12): (1022.384, 4.40e-7, 1.21e-6),
}


def grid(args):
    """The reference's published device grid on this GPU: its problems exactly (N = 10^5,
    10^6, 10^7 points uniform in [0, 1)^3 and charges uniform in [-1, 1) from SplitMix64
    0x00be9c06, then 1,000 sampled targets, fmm-bench/src/measure.rs:35-67; sources =
    targets, gradients, max level 16, <= 64 points per leaf, no near rule as in those runs,
    dense M2L), the mixed path beside the reference's f32 rows (p = 3, 6, 8) and the plain
    f64 path beside its f64 rows (p = 3, 6, 8, 12; its p = 18 is beyond this path's
    p <= 15).  One tuning cache for the grid, as the reference shares one per precision
    over its N.  Times: the median of the timed evaluations with charges and outputs in
    device memory, and with them in host memory (upload and download included, as the
    reference's rows).  Prints one JSON line with every row."""
    import tempfile

    import torch

    import fmm_amd
    from fmm_amd import points as P, tune

    dev = torch.device("cuda:0")
    tune_dir = args.tuning_cache or tempfile.mkdtemp(prefix="fmm_amd_tune")
    rows = []
    for n in (100_000, 1_000_000, 10_000_000):
        t0 = time.time()
        pts, q, sample = P.bench_cube(n, 1000, charges="uniform")
        plan = fmm_amd.Plan(pts, leaf_points=64)
        ep, eg = fmm_amd.direct_sum(pts, q, pts[sample])
        ep, eg = ep / (4 * np.pi), eg / (4 * np.pi)
        log(f"grid N={n}: {plan.nleaves} leaves on {plan.nlevels} levels, plan and direct sum in "
            f"{time.time() - t0:.1f}s")
        q_dev = torch.from_numpy(q).to(dev)
        phi = torch.zeros(n, dtype=torch.float64, device=dev)
        grad = torch.zeros(3 * n, dtype=torch.float64, device=dev)
        for precision, ref_prec, degrees in (("mixed", "f32", (3, 6, 8)), ("f64", "f64", (3, 6, 8, 12))):
            for p in degrees:
                tuning = None
                if args.tune_budget > 0:
                    tuning, _ = tune.resolve(plan, p, precision, True, 0, "dense", tune_dir, args.tune_budget)
                op = fmm_amd.Operator(plan, p, precision, True, 0, "dense", tuning)
                fmm = fmm_amd.Fmm(plan, op)

                def step():
                    op.load_charges(q_dev)
                    op.reset()
                    op.evaluate()
                    op.read_output_device(phi, grad)

                for _ in range(args.warmup):
                    step()
                torch.cuda.synchronize()
                dt = []
                for _ in range(args.steps):
                    torch.cuda.synchronize()
                    t = time.perf_counter()
                    step()
                    torch.cuda.synchronize()
                    dt.append(time.perf_counter() - t)
                ph = phi.cpu().numpy()
                gr = grad.cpu().numpy().reshape(-1, 3)
                e1 = float(np.linalg.norm(ph[sample] - ep) / np.linalg.norm(ep))
                e2 = float(np.linalg.norm(gr[sample] - eg) / np.linalg.norm(eg))
                ht = []
                for _ in range(max(3, args.steps // 2)):
                    t = time.perf_counter()
                    fmm.evaluate(q)
                    ht.append(time.perf_counter() - t)
                ms, hms = float(np.median(dt)) * 1e3, float(np.median(ht)) * 1e3
                pub = PUBLISHED_GRID.get((ref_prec, n, p))
                row = {"precision": precision, "reference_precision": ref_prec, "n": n, "p": p,
                       "ms": ms, "host_ms": hms, "mpts_s": n / ms / 1e3, "err_phi": e1, "err_grad": e2,
                       "published_ms": None if pub is None else pub[0],
                       "published_err_phi": None if pub is None else pub[1],
                       "published_err_grad": None if pub is None else pub[2],
                       "host_speedup_vs_published": None if pub is None else pub[0] / hms}
                rows.append(row)
                log(f"  {precision} p={p}: {ms:.3f} ms device-resident, {hms:.3f} ms host round trip "
                    f"(published {ref_prec} {pub[0] if pub else None} ms), errors {e1:.3g} / {e2:.3g} "
                    f"(published {pub[1] if pub else None} / {pub[2] if pub else None})")
                del fmm, op
        del plan
    line = {"metric": "the reference's published device grid on one MI355X (ms per evaluation)",
            "value": None, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "mixed: f32 P2P + bf16x3 M2L + f64 other; f64: every operator in f64", "data": "synthetic",
            "config": {"workload": "reference-grid", "leaf_points": 64, "max_level": 16, "near_rule": None,
                       "charges": "uniform [-1, 1)", "sampled_targets": 1000, "m2l": "dense",
                       "reference": "fmm-bench/results/phase4s-gh200.md:89-113 (H100 of a GH200)"},
            "rows": rows}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
