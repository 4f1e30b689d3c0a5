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

# algorithmic FP64 operations per cell (per point for the sampler) of each entry point, counted on
# the reference's generated kernels
FLOPS_PER_UNIT = {
    "seqvp_q1_soft_energy": 14445.,
    "seqvp_q1_barrier_energy": 8460.,
    "seqvp_q1_gradient": 12971. + 8093.,
    "seqvp_q1_ggn_diagonal": 27290.,
    "seqvp_jd_root": 4015.,
    "seqvp_spline_warp": 1078.,
}
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--inner", type=int, default=5)
    ap.add_argument("--shape", type=str, default="160,192,224")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("driver", "fixed"), default="driver")
    ap.add_argument("--fp32-screen", action="store_true", help="precision form: FP32 screening energies")
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


def cpu_baseline(shape_native, inner):
    """The NumPy oracle's cell sweeps of one outer step on a sub-volume, scaled by the cell count."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import jd_root as ojr, prescreen as ops, q1 as Q
    sub = (22, 26, 30)
    g = Q.Grid(sub, (1., 1., 1.))
    m = Q.Model(g, lam=2.8e-6, mu=.03, nu=.03, kappa=.002, delta=0., activation=.05)
    rng = np.random.default_rng(0)
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n) for n in sub), indexing="ij")
    u = np.stack([1.5 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(2 * np.pi * z)] * 3, axis=-1)
    ref = np.zeros_like(u)
    base = .9 * u
    b = rng.normal(size=u.shape)
    D = m.closed[3]
    if not ops.LIB.exists():
        from oracle import build as ob
        ob.build()
    with threadpool_limits(1):
        t0 = time.perf_counter()
        m.ggn_diagonal_nodes(m.ggn_diagonal_base(b), u)
        for _ in range(inner):
            m.gradient_nodes(np.zeros_like(u), u, ref)
            ojr.global_root(base, u - base, g, D, 0., 1.)
            ops.lower_bounds_c(u, g, D)
            m.soft_energy(u, ref, base)
            m.barrier_energy(u, base)
        dt = time.perf_counter() - t0
    cells_native = int(np.prod([n - 1 for n in shape_native]))
    return dict(value=round(dt * 1e3 * cells_native / g.n_cells, 1), unit="ms/outer-step", cores=1, kind="port",
                sample=f"NumPy oracle cell sweeps of one outer step ({inner} inner iterations: gradient, JD root, "
                       f"prescreen, soft+barrier energies; plus the GGN diagonal) on a {sub[0]}x{sub[1]}x{sub[2]} "
                       f"sub-volume ({g.n_cells} cells, {dt:.1f} s), scaled by the native cell count")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (a pair is not sharded "
              f"across GPUs here; independent pairs would be placed per GPU)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.spline import DeviceSpline, build_image_sampler
    from seq_vp_3d_amd.synthetic import CONFIG_NATIVE, PairSpec, make_pair

    dev = torch.device("cuda:0")
    shape = tuple(int(v) for v in args.shape.split(","))
    spec = CONFIG_NATIVE if shape == CONFIG_NATIVE.shape else PairSpec(shape, CONFIG_NATIVE.blobs, CONFIG_NATIVE.sigma,
                                                                       CONFIG_NATIVE.modes, CONFIG_NATIVE.jmin)

    def sampler(image, grid, points):
        sp = DeviceSpline(build_image_sampler(image.cpu().numpy(), grid), grid, dev)
        return sp.sample(points, with_gradient=False)[0]

    t0 = time.perf_counter()
    pair = make_pair(spec, dev, sampler)
    problem = Q1Problem(pair.grid, native_weights(shape), 0., .05)
    st = PairState(problem, pair.fixed, pair.moving, dev)
    theta0 = torch.zeros(st.nfree, dtype=torch.float64, device=dev)
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
        inner_cfg = PGConfig(max_iterations=args.inner, fp32_screen=args.fp32_screen)

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
    run(args.warmup, "warmup")
    torch.cuda.synchronize()
    _lib.TIMER = _lib.CallTimer()
    stats = dict(trials=0, accepted=0)
    t0 = time.perf_counter()
    run(args.steps, "step")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timer, _lib.TIMER = _lib.TIMER, None
    kern = timer.summary()
    ms_per_step = dt * 1e3 / args.steps
    ncell = int(np.prod([n - 1 for n in shape]))
    nnode = int(np.prod(shape))
    totals = {k: c * ms for k, (c, ms) in kern.items()}
    dominant = max((k for k in totals if k in FLOPS_PER_UNIT), key=lambda k: totals[k])
    units = nnode if dominant == "seqvp_spline_warp" else ncell
    calls, mean_ms = kern[dominant]
    achieved = FLOPS_PER_UNIT[dominant] * units / (mean_ms * 1e-3) / 1e12
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
                    shape_zyx=list(shape), pair="synthetic native recipe seed 1", trials=stats["trials"],
                    accepted=stats["accepted"]),
        roofline=dict(bound="mfma", achieved=round(achieved, 2), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                      frac=round(achieved / FP64_PEAK_TFLOPS, 4), traffic=None, kernel=dominant,
                      mean_ms=round(mean_ms, 4), note="FP64 vector peak (equal to the FP64 matrix peak on MI355X); "
                                                        "achieved = algorithmic flops per launch / mean launch time"),
        kernels={k: dict(calls=c, mean_ms=round(ms, 4)) for k, (c, ms) in sorted(kern.items())},
    )
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(shape, args.inner)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
