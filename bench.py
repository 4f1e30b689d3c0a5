"""Benchmark: ms per outer step of the registration hot path at 160x192x224 on one MI355X.

Workload (synthetic pair, built as the 160x192x224 native configuration: 200 Gaussian blobs, sigma
3-12 voxels, 6 sinusoid modes per component, true JD >= 0.5, noise 0.01): one outer step =
the metric build (GGN diagonal, image shift, T62 node JD, equilibration scale) followed by
`--inner` projected-gradient iterations, each with the full gradient, the equilibrated DST-I
direction, the T66 JD-root step bound, one trial (candidate map, whole-cell JD certificate of every cell,
spline sampling, energies and stable differences against the accepted state) and the promotion of
an accepted trial. Inputs are resident in HBM before the timed region; only scalars cross PCIe.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# algorithmic FP64 operations per cell (per point for the sampler) of each entry point: the + - * (an FMA as two)
# of the reference's generated Triton kernels for the same work
FLOPS_PER_UNIT = {
    "seqvp_q1_soft_energy": 14445.,
    "seqvp_q1_barrier_energy": 8460.,
    "seqvp_q1_trial_energy": 14445. + 8460.,
    "seqvp_q1_trial_energy_base_min": 14445. + 8460.,
    "seqvp_q1_gradient": 12971. + 8093.,
    "seqvp_q1_gradient_free": 12971. + 8093.,
    "seqvp_q1_ggn_diagonal": 27290.,
    "seqvp_jd_root": 4015.,
    "seqvp_spline_warp": 1078.,
}
# FP64 operations the build's kernels actually execute per cell (per node for the sampler): SQ_INSTS_VALU_FLOPS_FP64
# x 64 lanes per dispatch / units (rocprofv3 --pmc, profiles/valu_issue_native_pmc.md). The edge-difference
# geometry, the tabulated blend weights and the compile-time closed points do the reference's work with a third to a
# half of its operation count.
EXECUTED_FLOPS_PER_UNIT = {
    "seqvp_q1_soft_energy": 7223.,
    "seqvp_q1_barrier_energy": 1620.,
    "seqvp_q1_trial_energy": 8699.,
    "seqvp_q1_trial_energy_base_min": 7614.,
    "seqvp_q1_gradient": 7129. + 24.,
    "seqvp_q1_gradient_free": 7127. + 38.,
    "seqvp_q1_ggn_diagonal": 8781. + 38.,
    "seqvp_jd_root": 532. + 38.,
    "seqvp_spline_warp": 1015.,
}
FP64_PEAK_TFLOPS = 78.6
# FP64 vector FMAs one MI355X sustains on a register-only stream (bench.py --ceilings, profiles/fp64_ceilings_mi355x.md;
# about 2.25 GHz under load against the 2.4 GHz of the published peak)
FP64_VALU_SUSTAINED_TFLOPS = 66.9
# HBM bytes per launch at 160x192x224 from the FETCH_SIZE and WRITE_SIZE counter runs (rocprofv3 --pmc, one counter
# per run, profiles/hbm_traffic_native_pmc.md), each counter scaled by its reading for a copy of known size
# (--copy-probe: fetch x2.0, write x1.0)
TRAFFIC_BYTES_NATIVE = {
    "seqvp_q1_soft_energy": 986e6,
    "seqvp_q1_barrier_energy": 362e6,
    "seqvp_q1_trial_energy": 988e6,
    "seqvp_q1_trial_energy_base_min": 991e6,
    "seqvp_q1_gradient": 3666e6,
    "seqvp_q1_gradient_free": 3940e6,
    "seqvp_q1_joint_free": 5973e6,
    "seqvp_jd_root": 380e6,
    "seqvp_jd_root_free": 380e6,
    "seqvp_certificate_screen": 283e6,
    "seqvp_spline_warp": 610e6,
    "seqvp_dst_poisson": 2531e6,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--inner", type=int, default=5)
    ap.add_argument("--shape", type=str, default="160,192,224")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="driver workload: skip the batching-form measurement added to the line after the timed "
                         "region")
    ap.add_argument("--workload", choices=("driver", "fixed", "precision", "serve", "batch", "file"),
                    default="driver")
    ap.add_argument("--requests", type=int, default=24, help="serve workload: number of requests")
    ap.add_argument("--rate", type=float, default=2., help="serve workload: Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=4, help="serve/batch workloads: concurrent requests")
    ap.add_argument("--batched", action="store_true",
                    help="serve/batch workloads: also run the requests with their trial-energy sweeps in multi-pair "
                         "launches (batch.BatchGroup, launch shapes tuned online)")
    ap.add_argument("--max-wait", type=float, default=.002,
                    help="--batched: seconds the first pair of a launch waits for the others")
    ap.add_argument("--graphs", choices=("on", "off"), default="off",
                    help="trial and direction segments as captured HIP graphs (on) or issued call by call (off)")
    ap.add_argument("--start", choices=("zero", "half"), default="zero",
                    help="initial map: u = 0 (the configuration's start) or half the known displacement")
    ap.add_argument("--fp32-screen", action="store_true", help="precision form: FP32 screening energies")
    ap.add_argument("--outer", type=int, default=20, help="precision workload: outer steps of each FP64 solve")
    ap.add_argument("--busy", nargs=2, metavar=("KERNEL_TRACE_CSV", "FORMS_LINE_FILE"),
                    help="GPU busy share of the serial and served phases of a batch/serve run traced by rocprofv3 "
                         "--kernel-trace; prints one JSON line and exits")
    ap.add_argument("--ceilings", action="store_true",
                    help="measure the FP64 MFMA and FP64 VALU rates one GPU sustains on register-only instruction "
                         "streams (seqvp_fp64_ceiling_probe), print one JSON line and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="before the workload, one seqvp_scale launch over 1 GiB (reads 1 GiB, writes 1 GiB): the "
                         "plain copy of known size that calibrates the FETCH_SIZE/WRITE_SIZE counters")
    return ap.parse_args()


def outer_step(st, inner, stats):
    st.metric()
    for _ in range(inner):
        st.gradient()
        st.direction()
        (slope,) = st.dots((st.g, st.p))
        tau0 = 1.
        root, invalid, _ = st.jd_bound(tau0)
        tau = tau0 if root == float("inf") else min(tau0, .9 * root * tau0)
        for _attempt in range(4):
            tr = st.trial(tau)
            ok = tr.ok and tr.unproven_cells == 0 and -tr.total_difference >= 1e-4 * tau * abs(slope)
            stats["trials"] += 1
            if ok:
                st.promote(tr.energy)
                stats["accepted"] += 1
                break
            tau *= .5


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(shape_native, inner, sub=(20, 22, 24)):
    """One full certified outer step of the oracle's restatement of the reference's CPU path (oracle/cpu_step.py,
    single-threaded NumPy), run on independent sub-volumes, one process per host core at once; the native
    ms/outer-step is the wall time scaled by native cells / cells processed."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    from oracle.cpu_step import outer_step_seconds
    cores = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    t0 = time.perf_counter()
    with ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as pool:
        runs = list(pool.map(outer_step_seconds, [sub] * cores, range(cores), [inner] * cores))
    wall = time.perf_counter() - t0
    step = max(r[0] for r in runs)                     # the slowest worker's outer step (setup excluded)
    cells = sum(r[1] for r in runs)
    cells_native = 1
    for n in shape_native:
        cells_native *= n - 1
    return dict(value=round(step * 1e3 * cells_native / cells, 1), unit="ms/outer-step", cores=cores, kind="port",
                cpu_model=_cpu_model(), nproc=os.cpu_count(),
                sample=(f"one certified outer step of the oracle's restatement of the reference CPU path (T73 metric, "
                        f"{inner} PG steps with gradient, DST-I direction, BB2, witness + T66 predictions, fully "
                        f"certified trials, spline sampling, stable E/F comparisons, outer norms) on {cores} "
                        f"independent {sub[0]}x{sub[1]}x{sub[2]} synthetic pairs at once, one single-threaded NumPy "
                        f"process per core ({cells} cells, slowest step {step:.1f} s, pool wall {wall:.1f} s), "
                        f"scaled by the native cell count; outer reasons {sorted(set(r[2] for r in runs))}"))


def precision_form(args, pair, problem, dev):
    """The precision form at the pair's shape: FP32 8/4/2/1 pyramid pre-optimisation (native flags), certified
    transfer and `--outer` certified FP64 outer steps, beside `--outer` FP64 outer steps from u = 0. One JSON line
    with the per-level FP32-vs-FP64 errors and the final energies of both routes."""
    import numpy as np
    import torch

    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.preopt import PreoptConfig, preoptimized_solve
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.spline import DeviceSpline

    spline = DeviceSpline.from_image(pair.moving, pair.grid)
    cfg = PreoptConfig.native()

    def log(label):
        return lambda row: print(f"{label} outer {row['outer_i']} {row['reason']} E {row['E_candidate']} "
                                 f"{row['seconds'] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # warm the kernels once (a short pre-optimisation on the same pair)
    preoptimized_solve(problem, pair.fixed, pair.moving, dev, PreoptConfig(iterations=(2, 2, 2, 2)), max_outer=1,
                       sampler=spline, report_errors=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, res, rec = preoptimized_solve(problem, pair.fixed, pair.moving, dev, cfg, max_outer=args.outer,
                                      observer=log("preopt-route"), sampler=spline,
                                      inner_config=PGConfig(jd_step_global=True))
    torch.cuda.synchronize()
    route_s = time.perf_counter() - t0
    if res is None:
        raise SystemExit(f"no certified transfer: {rec['transfer']}")
    st64 = PairState(problem, pair.fixed, pair.moving, dev, sampler=spline)
    st64.set_state(theta=torch.zeros(st64.nfree, dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res64 = solve_sequential(st64, SequentialConfig(), PGConfig(jd_step_global=True), observer=log("fp64-only"),
                             max_outer=args.outer)
    torch.cuda.synchronize()
    fp64_s = time.perf_counter() - t1
    levels = [dict(level=lv["level"], shape_zyx=lv["shape_zyx"], iterations=lv["iterations"],
                   seconds=round(lv["seconds"], 4), graph=lv["graph"], first_total=lv["first"]["total"],
                   last_total=lv["last"]["total"], start_errors=lv["start_errors"], end_errors=lv["end_errors"])
              for lv in rec["preopt"]["levels"]]
    e_pre, e_64 = res.energy.total, res64.energy.total
    line = dict(metric="precision form @%s" % "x".join(map(str, pair.grid.shape)), unit="s", dtype="f32+f64",
                data="synthetic", config=dict(workload="FP32 8/4/2/1 pyramid pre-optimisation (native flags) + "
                                                        "certified transfer + FP64 certified outer steps",
                                              shape_zyx=list(pair.grid.shape), outer_steps=args.outer,
                                              preopt=dict(iterations=list(cfg.iterations), lr=cfg.lr,
                                                          sigma_field=cfg.sigma_field, mu=list(cfg.mu))),
                preopt_seconds=round(rec["preopt_seconds"], 4), transfer_alpha=rec["transfer"]["alpha"],
                transfer_seconds=round(rec["transfer"]["seconds"], 4), handoff_energy=rec["handoff_energy"],
                levels=levels, final_energy_preopt_route=e_pre, final_energy_fp64_only=e_64,
                final_energy_ratio=e_pre / e_64, route_seconds=round(route_s, 3), fp64_only_seconds=round(fp64_s, 3),
                outer_status=dict(preopt_route=res.status, fp64_only=res64.status))
    print(json.dumps(line), flush=True)


SERVE_SIZES = ((96, 96, 96), (128, 128, 128), (160, 192, 144), (160, 192, 224))
BATCH_SIZES = ((96, 96, 96),) * 8 + ((128, 128, 128),) * 8 + ((160, 192, 144),) * 4 + ((160, 192, 224),) * 4


def file_form(args, pair, dev):
    """The pair-folder form: the synthetic pair is written as pair.npz + manifest.json (untimed), then run_pair --
    hash check and load, upload, certified FP64 solve of `--steps` outer steps from u = 0, the final state's
    certificate, final_native.npz/.json, summary.json, history.jsonl -- is timed end to end. The independent CPU
    re-certification of the stored map is timed separately (it is an audit after the run)."""
    import shutil
    import tempfile
    from pathlib import Path

    import torch

    from seq_vp_3d_amd import pairio

    root = Path(tempfile.mkdtemp(prefix="seqvp_pair_"))
    try:
        folder = root / "pair"
        t0 = time.perf_counter()
        pairio.write_pair(folder, pair.fixed.cpu().numpy(), pair.moving.cpu().numpy(), pair="synthetic_native")
        print(f"pair folder written in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        pairio.run_pair(folder, root / "warm", device=dev, route="native", max_outer=max(1, args.warmup))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        summary = pairio.run_pair(folder, root / "run", device=dev, route="native", max_outer=args.steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        t1 = time.perf_counter()
        cert = pairio.recertify(root / "run", folder, workers=8)
        audit = time.perf_counter() - t1
        outer = max(1, summary["outer_rows"])
        line = dict(metric="ms/outer-step @%s" % "x".join(map(str, pair.grid.shape)), unit="ms/outer-step",
                    value=round(wall * 1e3 / outer, 3), higher_is_better=False, n_gpus=1, steps=args.steps,
                    warmup=args.warmup, dtype="f64", data="synthetic",
                    config=dict(workload="pair folder -> run folder: load + certified FP64 solve from u = 0 + "
                                         "final certificate + export, end to end", shape=list(pair.grid.shape)),
                    status=summary["status"], outer_steps=summary["outer_rows"],
                    native_commits=summary["native_commits"], wall_seconds=round(wall, 3),
                    e2e_seconds=round(summary["e2e_seconds"], 3), input_seconds=round(summary["input_seconds"], 3),
                    solve_seconds=round(summary["solve_call_wall_seconds"], 3),
                    solve_ms_per_outer_step=round(summary["solve_call_wall_seconds"] * 1e3 / outer, 3),
                    final_export_seconds=round(summary["final_export_seconds"], 3),
                    stored_array_io_seconds=round(summary["stored_array_io_seconds"], 3),
                    recertify_seconds=round(audit, 3), recertified=cert.status,
                    recertified_min_jd_lower_bound=cert.min_jd_lower_bound, mse_ratio=summary["ratio"])
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


def gpu_busy(trace_csv, line_file):
    """GPU busy share of the serial and served phases of a forms run: the union of the kernel intervals of a
    rocprofv3 kernel trace (CSV) inside each phase's CLOCK_MONOTONIC window from the run's JSON line."""
    import csv
    line = json.loads([ln for ln in open(line_file) if ln.startswith("{")][-1])
    spans = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                   for r in csv.DictReader(open(trace_csv)))
    out = {}
    for phase, (t0, t1) in line["phase_monotonic_ns"].items():
        sel = [(a, b) for a, b, _ in spans if a >= t0 and b <= t1]
        busy, cur = 0, None
        for a, b in sel:
            if cur is None or a > cur[1]:
                busy += 0 if cur is None else cur[1] - cur[0]
                cur = [a, b]
            else:
                cur[1] = max(cur[1], b)
        busy += 0 if cur is None else cur[1] - cur[0]
        out[phase] = dict(wall_s=round((t1 - t0) / 1e9, 3), kernels=len(sel),
                          kernel_time_s=round(sum(b - a for a, b in sel) / 1e9, 3),
                          busy_s=round(busy / 1e9, 3), busy_share=round(busy / (t1 - t0), 3))
    print(json.dumps(dict(metric="GPU busy share per phase", workload=line["config"]["workload"], **out)), flush=True)


def fp64_ceilings(dev):
    """Sustained FP64 rates on register-only streams, timed with HIP events on the launch stream: the best over
    1-8 workgroups per CU (256 CUs) of v_mfma_f64_16x16x4f64 and of vector FMAs, against the published 78.6 TFLOP/s."""
    import torch

    from seq_vp_3d_amd import _lib
    out = torch.empty(8 * 256 * 256, dtype=torch.float64, device=dev)
    stream = _lib.stream_handle(dev)
    rows = {}
    for kind, name, iters, per in ((0, "mfma", 2000, lambda b, it: b * 4 * it * 8 * 2048.),
                                   (1, "valu", 20000, lambda b, it: b * 256 * it * 8 * 2.)):
        best = []
        for per_cu in (1, 2, 4, 8):
            blocks = 256 * per_cu
            _lib.call("seqvp_fp64_ceiling_probe", kind, blocks, 10, _lib.ptr(out), stream)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            _lib.call("seqvp_fp64_ceiling_probe", kind, blocks, iters, _lib.ptr(out), stream)
            b.record()
            b.synchronize()
            tf = per(blocks, iters) / (a.elapsed_time(b) * 1e-3) / 1e12
            best.append((round(tf, 2), per_cu))
        rows[name] = dict(tflops=max(best)[0], workgroups_per_cu=max(best)[1], sweep=best)
    print(json.dumps(dict(metric="sustained FP64 rate on register-only streams", unit="TFLOP/s",
                          published_peak=FP64_PEAK_TFLOPS, **rows)), flush=True)


def serving_form(args, dev):
    """The batching (all requests at t = 0) and serving (Poisson arrivals) forms against serial one-pair-at-a-time
    execution of the same requests; each request runs `--outer` certified outer steps from u = 0 and its output is
    re-certified on the CPU. One JSON line: certified pairs/hour of both, p50/p99 latency, and whether every
    result is bitwise the same alone and in the mix."""
    print(json.dumps(forms_line(args, dev)), flush=True)


def forms_line(args, dev):
    """The measurement of serving_form as a dict."""
    import numpy as np
    import torch

    from seq_vp_3d_amd.problem import make_grid_3d
    from seq_vp_3d_amd.serving import Request, latency_summary, poisson_offsets, serial, serve
    from seq_vp_3d_amd.synthetic import device_min_jd, device_sampler, make_pair, scaled_spec

    if args.workload == "batch":
        shapes, seeds = BATCH_SIZES, [1000 + i for i in range(len(BATCH_SIZES))]
    else:
        pick = np.random.default_rng(7).integers(0, len(SERVE_SIZES), size=args.requests)
        shapes, seeds = [SERVE_SIZES[k] for k in pick], [2000 + i for i in range(args.requests)]
    t0 = time.perf_counter()
    reqs = []
    for i, (shape, seed) in enumerate(zip(shapes, seeds)):
        grid = make_grid_3d(nx=shape[2], ny=shape[1], nz=shape[0], hx=1., hy=1., hz=1.)
        pair = make_pair(scaled_spec(shape, seed), dev, device_sampler(dev), device_min_jd(grid, dev))
        reqs.append(Request(pair.fixed.cpu().pin_memory(), pair.moving.cpu().pin_memory(), max_outer=args.outer,
                            request_id=i))
        del pair
    torch.cuda.synchronize()
    print(f"generated {len(reqs)} pairs in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # warm-up: every shape once (kernels, allocator pools, host caches)
    warm = {}
    for r in reqs:
        warm.setdefault(r.shape, Request(r.fixed, r.moving, max_outer=1, request_id=-1))
    serial(list(warm.values()), dev)
    torch.cuda.synchronize()
    phases = {}                     # CLOCK_MONOTONIC ns, the clock of rocprofv3's kernel timestamps
    t_a = time.monotonic_ns()
    alone, serial_s = serial(reqs, dev)
    torch.cuda.synchronize()
    phases["serial"] = [t_a, time.monotonic_ns()]
    print(f"serial {serial_s:.2f}s", file=sys.stderr, flush=True)
    offsets = None if args.workload == "batch" else poisson_offsets(len(reqs), args.rate, 7)
    t_a = time.monotonic_ns()
    mixed, wall, server = serve(reqs, offsets, dev, slots=args.slots)
    torch.cuda.synchronize()
    phases["served"] = [t_a, time.monotonic_ns()]
    print(f"served {wall:.2f}s", file=sys.stderr, flush=True)
    batched = None
    if args.batched:
        # the same requests again with every trial-energy sweep in the BatchGroup's multi-pair launches (tuned
        # online: the first launches of each size class try the candidate shapes)
        from seq_vp_3d_amd.batch import BatchGroup
        group = BatchGroup(dev, max_wait=args.max_wait)
        t_a = time.monotonic_ns()
        bres, bwall, _ = serve(reqs, offsets, dev, slots=args.slots, batch=group)
        torch.cuda.synchronize()
        phases["batched"] = [t_a, time.monotonic_ns()]
        print(f"batched {bwall:.2f}s, {group.launches} launches for {group.pairs_launched} pair sweeps",
              file=sys.stderr, flush=True)
        batched = dict(pairs_per_hour=round(len(reqs) / bwall * 3600, 1), wall_seconds=round(bwall, 3),
                       speedup_vs_serial=round(serial_s / bwall, 3), speedup_vs_streams=round(wall / bwall, 3),
                       launches=group.launches, pairs_per_launch=round(group.pairs_launched / max(group.launches, 1), 2),
                       max_wait_s=args.max_wait,
                       same_as_alone=all(a.theta_digest == b.theta_digest and a.energy == b.energy
                                         for a, b in zip(alone, bres)),
                       recertified=f"{sum(r.recertified == 'certified_feasible' for r in bres)}/{len(bres)}",
                       latency_seconds={k: round(v, 4) for k, v in latency_summary(bres).items()},
                       tuner={str(k): v for k, v in group.tuner.report().items()})
    same = all(a.theta_digest == b.theta_digest and a.energy == b.energy for a, b in zip(alone, mixed))
    for a, b in zip(alone, mixed):
        if a.theta_digest != b.theta_digest:
            k = next((i for i, (x, y) in enumerate(zip(a.trajectory, b.trajectory)) if x != y),
                     min(len(a.trajectory), len(b.trajectory)))
            print(f"request {a.request_id} {a.shape}: attempts {len(a.trajectory)} alone / {len(b.trajectory)} "
                  f"mixed, first difference at attempt {k}: alone {a.trajectory[k - 1:k + 2]} mixed "
                  f"{b.trajectory[k - 1:k + 2]}", file=sys.stderr, flush=True)
    # final-energy tolerance of the forms: each served result against the same request run alone on the plain path
    ediff = max(abs(a.energy - b.energy) / max(abs(a.energy), 1e-300) for a, b in zip(alone, mixed))
    statuses = {}
    for r in mixed:
        statuses[r.status] = statuses.get(r.status, 0) + 1
    recert = sum(r.recertified == "certified_feasible" for r in mixed)
    lat_mix, lat_alone = latency_summary(mixed), latency_summary(alone)
    n = len(reqs)
    line = dict(metric="certified pairs/hour & p50/p99 latency served", unit="pairs/hour",
                value=round(n / wall * 3600, 1), higher_is_better=True, n_gpus=1, dtype="f64", data="synthetic",
                config=dict(workload=("batching: %d pairs of uneven sizes submitted at once" % n
                                      if args.workload == "batch" else
                                      "serving: %d requests, Poisson arrivals at %.2f/s" % (n, args.rate)),
                            sizes={"x".join(map(str, k)): sum(1 for r in reqs if r.shape == k)
                                   for k in sorted(set(r.shape for r in reqs))},
                            outer_steps_per_request=args.outer, slots=args.slots,
                            pair_seeds=f"{seeds[0]}..{seeds[-1]}",
                            result=("each request: the certified FP64 PG solve from u = 0 until its outer loop stops "
                                    f"(convergence, or {args.outer} accepted outer steps), T66 root on; the output "
                                    "map re-certified by the CPU interval certificate before it is returned")),
                final_energy_max_rel_diff_vs_alone=ediff, final_status=statuses,
                serial_pairs_per_hour=round(n / serial_s * 3600, 1), speedup_vs_serial=round(serial_s / wall, 3),
                wall_seconds=round(wall, 3), serial_seconds=round(serial_s, 3),
                latency_seconds=dict(mixed={k: round(v, 4) for k, v in lat_mix.items()},
                                     alone={k: round(v, 4) for k, v in lat_alone.items()}),
                max_concurrent=max(a[4] for a in server.admissions), recertified=f"{recert}/{n}",
                same_alone_and_in_mix=same, phase_monotonic_ns=phases, batched_launches=batched,
                per_request=[dict(id=r.request_id, shape=list(r.shape), latency=round(r.latency, 4),
                                  solve=round(r.seconds["solve"], 4), setup=round(r.seconds["setup"], 4),
                                  recertify=round(r.seconds["recertify"], 4), status=r.status) for r in mixed])
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (a pair is not sharded "
              f"across GPUs here; independent pairs would be placed per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.busy:
        return gpu_busy(*args.busy)
    import numpy as np
    import torch

    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.problem import make_grid_3d
    from seq_vp_3d_amd.synthetic import CONFIG_NATIVE, PairSpec, device_min_jd, device_sampler, make_pair

    dev = torch.device("cuda:0")
    if args.ceilings:
        return fp64_ceilings(dev)
    if args.copy_probe:
        n = 1 << 27                                   # 2^27 doubles = 1 GiB
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        _lib.call("seqvp_scale", _lib.ptr(a), 2., _lib.ptr(b), n, _lib.stream_handle(dev))
        torch.cuda.synchronize()
        del a, b
    if args.workload in ("serve", "batch"):
        return serving_form(args, dev)
    shape = tuple(int(v) for v in args.shape.split(","))
    spec = CONFIG_NATIVE if shape == CONFIG_NATIVE.shape else PairSpec(shape, CONFIG_NATIVE.blobs, CONFIG_NATIVE.sigma,
                                                                       CONFIG_NATIVE.modes, CONFIG_NATIVE.jmin)

    t0 = time.perf_counter()
    native_grid = make_grid_3d(nx=shape[2], ny=shape[1], nz=shape[0], hx=1., hy=1., hz=1.)
    pair = make_pair(spec, dev, device_sampler(dev), device_min_jd(native_grid, dev))
    problem = Q1Problem(pair.grid, native_weights(shape), 0., .05)
    if args.workload == "precision":
        return precision_form(args, pair, problem, dev)
    if args.workload == "file":
        return file_form(args, pair, dev)
    st = PairState(problem, pair.fixed, pair.moving, dev)
    theta0 = torch.zeros(st.nfree, dtype=torch.float64, device=dev)
    if args.start == "half":
        st.restrict((.5 * pair.displacement).contiguous(), theta0)
    st.set_state(theta=theta0)
    setup_s = time.perf_counter() - t0
    print(f"setup {setup_s:.1f}s amplitude {pair.amplitude:.3f} E0 {st.energy.total:.6e}", file=sys.stderr, flush=True)
    stats = dict(trials=0, accepted=0)
    if args.workload == "fixed":
        def run(n, label):
            for k in range(n):
                outer_step(st, args.inner, stats)
                print(f"{label} {k} done E {st.energy.total:.6e}", file=sys.stderr, flush=True)
    else:
        from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
        # the certified solver with the cubic JD-root bound on the device (T66, jd_step_global)
        inner_cfg = PGConfig(max_iterations=args.inner, fp32_screen=args.fp32_screen, jd_step_global=True)

        def run(n, label):
            def obs(row):
                stats["trials"] += row["inner_trials"]
                stats["accepted"] += row["inner_accepted_steps"]
                print(f"{label} outer {row['outer_i']} {row['reason']} inner {row['inner_status']} "
                      f"steps {row['inner_accepted_steps']} trials {row['inner_trials']} E {row['E_candidate']} "
                      f"GF {row['GF_candidate']} {row['seconds']*1e3:.1f} ms", file=sys.stderr, flush=True)
            res = solve_sequential(st, SequentialConfig(), inner_cfg, observer=obs, max_outer=n)
            stats["status"] = res.status
            if res.outer_accepted < n:
                raise SystemExit(f"outer loop stopped early: {res.status} {res.reason}")
    if args.graphs == "off":
        st.graphs = None
    run(args.warmup, "warmup")
    torch.cuda.synchronize()
    if st.graphs is None:
        # HIP events around the calls of the entry points that can be the dominant kernel (every bracket costs
        # host time, which the timed region would include)
        _lib.TIMER = _lib.CallTimer(names=FLOPS_PER_UNIT)
    stats = dict(trials=0, accepted=0)
    t0 = time.perf_counter()
    run(args.steps, "step")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timed_stats = dict(stats)
    timed, _lib.TIMER = _lib.TIMER, None
    # The means of every entry point: a second pass of as many outer steps, after the timed region, with every call
    # issued one by one and bracketed (captured segments run as graph launches, which events cannot bracket per
    # kernel); the timed region's own means replace the second pass's where it has them.
    graphs, st.graphs = st.graphs, None
    _lib.TIMER = _lib.CallTimer()
    run(args.steps, "timing")
    torch.cuda.synchronize()
    st.graphs = graphs
    stats = timed_stats
    timer, _lib.TIMER = _lib.TIMER, None
    kern = timer.summary()
    if timed is not None:
        kern.update(timed.summary())
    ms_per_step = dt * 1e3 / args.steps
    ncell = int(np.prod([n - 1 for n in shape]))
    nnode = int(np.prod(shape))
    totals = {k: c * ms for k, (c, ms) in kern.items()}
    dominant = max((k for k in totals if k in FLOPS_PER_UNIT), key=lambda k: totals[k])
    units = nnode if dominant == "seqvp_spline_warp" else ncell
    calls, mean_ms = kern[dominant]
    reference_rate = FLOPS_PER_UNIT[dominant] * units / (mean_ms * 1e-3) / 1e12
    executed = EXECUTED_FLOPS_PER_UNIT[dominant] * units / (mean_ms * 1e-3) / 1e12
    line = dict(
        metric="ms/outer-step @160x192x224", value=round(ms_per_step, 3), unit="ms/outer-step", n_gpus=1,
        steps=args.steps, warmup=args.warmup, ms_per_step=round(ms_per_step, 3), higher_is_better=False,
        scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload=("outer step of the certified FP64 PG solver (inner solve of up to "
                              f"{args.inner} accepted PG steps with BB2, T66 root bound, Armijo/resolution "
                              "backtracking, prescreen-certified trials; outer E/F comparison, commit, reference "
                              "refresh)") if args.workload == "driver" else
                             ("fixed outer-step sequence of the FP64 hot path (metric build + "
                              f"{args.inner} PG iterations, trial + prescreen + promotion each)"),
                    shape_zyx=list(shape), pair="synthetic native recipe seed 1", start=args.start,
                    trials=stats["trials"], accepted=stats["accepted"], graphs=args.graphs),
        roofline=dict(bound="valu", achieved=round(executed, 2), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                      frac=round(executed / FP64_PEAK_TFLOPS, 4),
                      sustained_valu_ceiling=FP64_VALU_SUSTAINED_TFLOPS,
                      frac_of_sustained=round(executed / FP64_VALU_SUSTAINED_TFLOPS, 4),
                      traffic=TRAFFIC_BYTES_NATIVE.get(dominant) if shape == (160, 192, 224) else None,
                      kernel=dominant, mean_ms=round(mean_ms, 4),
                      reference_count_rate=round(reference_rate, 2),
                      note="FP64 vector (VALU) roof: the cell sweeps are FP64 vector arithmetic, not MFMA and not "
                           "HBM-bound; achieved = the FP64 operations this kernel executes per launch "
                           "(SQ_INSTS_VALU_FLOPS_FP64 x 64 per cell, rocprofv3 --pmc) / its mean launch time (HIP "
                           "events on its stream); reference_count_rate = the reference's operator count for the "
                           "same work / the same time (above the peak because this build needs fewer operations)",
                      kernel_timing=("HIP events on the launch stream in the timed region (around the entry "
                                      "points that can be the dominant one; the other entry points' means in "
                                      "`kernels` from a second pass of as many outer steps after it)"
                                      if args.graphs == "off"
                                      else "HIP events on the launch stream in a second pass of as many outer steps "
                                           "with the captured segments issued call by call (graph launches cannot "
                                           "be bracketed per kernel)")),
        kernels={k: dict(calls=c, mean_ms=round(ms, 4)) for k, (c, ms) in sorted(kern.items())},
    )
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(shape, args.inner)
    if args.workload == "driver" and not args.no_forms and shape == CONFIG_NATIVE.shape:
        # the batching form (24 pairs of uneven sizes, each to a certified result after 10 outer steps) beside the
        # plain path, after the timed region: certified pairs/hour served, batched and serial, p50/p99 latency
        del st, pair
        torch.cuda.empty_cache()
        import types
        fargs = types.SimpleNamespace(workload="batch", outer=10, slots=args.slots, batched=True,
                                      max_wait=args.max_wait, requests=24, rate=args.rate)
        f = forms_line(fargs, dev)
        line["forms"] = dict(
            batch=dict(metric=f["metric"], served_pairs_per_hour=f["value"],
                       serial_pairs_per_hour=f["serial_pairs_per_hour"], speedup_vs_serial=f["speedup_vs_serial"],
                       batched_launch=f["batched_launches"], latency_seconds=f["latency_seconds"],
                       sizes=f["config"]["sizes"], outer_steps_per_request=10, recertified=f["recertified"],
                       same_alone_and_in_mix=f["same_alone_and_in_mix"],
                       final_energy_max_rel_diff_vs_alone=f["final_energy_max_rel_diff_vs_alone"]),
            serving="bench.py --workload serve --requests 200 --rate 2 (about two minutes of arrivals)")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
