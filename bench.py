#!/usr/bin/env python3
"""Benchmark of the APT time step on one MI355X.

Workload (default): 4096^2 cells, the nonlinear drop-in physics -- wavy grid
x = xi + 0.05 sin(2 pi xi) sin(2 pi eta) (same for y), kappa = T^2.5,
background T = 0.01 with a seeded hot disk T = 1 of radius 0.05, Dirichlet
T = 1 on the left edge, zero-flux elsewhere, adaptive dt from next_dt
(dt0 = 1e-6), residual tolerance 1e-8 (atol on the FP64 RMS residual,
rtol = 0), 10 steps.  A "step" is one time step: the full pseudo-transient
solve to tolerance.  The temperature stays in HBM for the whole timed region.
`--workload spots_3d` runs the 3D mixed-precision case instead: 256^3 cells,
x = xi + 0.04 sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta) (same for y, z),
kappa = T^3, 8 seeded Gaussian spots on T = 0.01, zero-flux faces, dt0 = 1e-5.

Timed in the same run, from the same state: the mixed-precision step (FP32
correction sweeps inside the FP64 defect-correction loop) with the learned
warm start attached (the committed network, consulted where the saving
expected from it exceeds a consultation's cost; a consulted step starts from
the network's prediction when that has the smaller FP64 residual, else from
T^n), the same step without the network (config.without_learned_warm_start),
the same step with the network consulted at every step from the second
(config.learned_consulted_every_step) and the plain FP64 APT.  `value` is the
s/time-step of the first, the method the metric names (FP32 + learned); in
3D, where the warm start does not apply, of the mixed-precision step.

Prints one JSON line: s/time-step (lower is better), with
  roofline      the dominant kernel's compulsory HBM bytes per launch (every
                array the launch needs read once and every array it updates
                written once: cells x bytes per cell per launch, whatever the
                number of pseudo-iterations the launch runs) / its mean launch
                duration (HIP events on the solver's stream), against 8 TB/s;
                `t_eff` is the same launch under the APT effective-throughput
                convention (A_eff = (2 D_u + D_k) bytes per cell and
                pseudo-iteration, charged once per iteration), `traffic` the
                HBM bytes per launch rocprofv3's counters measured (source and
                date in `traffic_source`)
  cpu_baseline  the oracle (CPU restatement of the reference APT, OpenMP on
                OMP_NUM_THREADS threads) on the first steps of the same
                workload, with the GPU's times for the same steps beside it;
                in 2D also the error of the mixed-precision step against the
                FP64 implicit-direct oracle on the 512^2 Gaussian subsample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# A_eff of the APT effective-throughput convention, bytes per cell and
# pseudo-iteration: (2 D_u + D_k) x bytes per value
BYTES_PER_CELL = {("fp64", 2): 96, ("fp32_dc", 2): 48, ("fp64", 3): 144, ("fp32_dc", 3): 72}
# compulsory bytes per cell and launch of the dominant kernels: what one launch
# must read (coefficients, sources, iterates) and write (iterates) once, for all
# the iterations it runs.
#   FP64 2D, two iterations: u, q_xi, q_eta read and written (48), u_prev and
#     g11, g22, c12, alpha_c, beta_c read (48)
#   FP32 2D correction sweep, eight iterations: e, q_xi, q_eta read and written
#     (24), five folded metric terms, P, M, W of both face sets, C and the
#     right-hand side read (52)
#   FP64 3D, one iteration: u and three fluxes read and written (64), u_prev and
#     nine coefficients read (80)
#   FP32 3D correction sweep, two iterations: e and three fluxes read and
#     written (32), nine folded metric terms, P, M, W of three face sets, C and
#     the right-hand side read (80)
COMPULSORY_BYTES_PER_CELL = {("fp64", 2): 96, ("fp32_dc", 2): 76, ("fp64", 3): 144, ("fp32_dc", 3): 112}


# workloads: mesh and T0 builder, boundary conditions (device, oracle), kappa
# law (kappa0, T exponent, rho exponent), dimension, default size and dt0
WORKLOADS = {
    "hot_disk_2d": dict(ndim=2, n=4096, dt0=1e-6, kappa=(1.0, 2.5, 0.0),
                        bcs=[("dirichlet", 1.0)] + [("neumann", 0.0)] * 3),
    "spots_3d": dict(ndim=3, n=256, dt0=1e-5, kappa=(1.0, 3.0, 0.0), bcs=[("neumann", 0.0)] * 6),
}


def make_workload(name, n, seed):
    if name == "spots_3d":
        from cases import gaussian_spots_3d

        return gaussian_spots_3d(n, seed=seed)
    from cases import hot_disk_2d

    return hot_disk_2d(n, seed=seed)


def device_bcs(spec):
    from curvilineardiffusionmanuscript_amd import solver as S

    return [S.DirichletBC(v) if kind == "dirichlet" else S.NeumannBC() for kind, v in spec]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(args, mesh, T0, method, net=None, tuning=None):
    """Warm-up steps, then args.steps timed steps of `method` from T0; with
    `net`, the learned warm start is attached (its back-off state reset with
    the solver's)."""
    from curvilineardiffusionmanuscript_amd import solver as S

    w = WORKLOADS[args.workload]
    law = S.PowerLawConductivity(*w["kappa"])
    sol = S.PseudoTransientSolver(mesh, device_bcs(w["bcs"]), tuning=tuning)
    kw = dict(rtol=0.0, atol=args.tol, max_iter=10**7, precision=method)

    def reset():
        sol.reset_state()
        if net is not None:
            net.attach(sol)
        S.update_conductivity(sol, mesh, T0, 1.0, 1.0, law)
        sol.set_temperature(T0)

    reset()
    dt = args.dt0
    for _ in range(args.warmup):
        st, dt = sol.step_resident(dt, None, 1.0, law, **kw)
    reset()
    dt = args.dt0
    iters, checks, steps, outer, warm, consulted, step_times = [], 0, [], [], 0, 0, []
    t0 = time.perf_counter()  # step_resident returns after the step's work has finished
    for k in range(args.steps):
        st, dt_next = sol.step_resident(dt, None, 1.0, law, **kw)
        iters.append(st.niter)
        outer.append(st.outer)
        checks += st.nchecks
        warm += int(st.warm_start)
        consulted += int(st.warm_consulted)
        steps.append((dt, st.niter, st.abs_err, st.time))
        step_times.append(st.time)
        dt = dt_next
    elapsed = time.perf_counter() - t0
    for k, (d, it, err, tm) in enumerate(steps[: min(4, len(steps))]):
        log(f"{method} step {k}: dt={d:.3e} niter={it} abs_err={err:.3e} time={tm * 1e3:.2f} ms")
    T_end = sol.get_temperature()
    mean_ms = float("nan")
    if args.sweep_samples > 0 and net is None:
        mean_ms = sol_bench(sol, steps[-1][0], method, args.sweep_samples)
    sol.close()
    return dict(elapsed=elapsed, iters=iters, outer=outer, checks=checks, T_end=T_end, sweep_ms=mean_ms,
                abs_err=[s[2] for s in steps], warm_used=warm, warm_consulted=consulted, step_times=step_times)


# dominant kernel of each method and dimension, with the pseudo-iterations
# one launch runs and the cdiff_bench_sweeps mode bits that launch it (bit 0:
# the residual-check variant -- every launch of the 2D correction solve
# checks; bit 1: two FP64 iterations per launch; bit 2: the correction sweep)
KERNELS = {
    ("fp64", 2): ("sweep2d_pair_kernel<double, false, 8>", 2, 0 | 2),
    ("fp32_dc", 2): ("sweep2d_lin_march_kernel<8, true, 1>", 8, 4 | 1),
    ("fp64", 3): ("sweep3d_col_kernel<double, false, true>", 1, 0),
    ("fp32_dc", 3): ("sweep3d_lin2_kernel<true>", 2, 4 | 1),
}


def sol_bench(sol, dt, method, n):
    import ctypes as C

    from curvilineardiffusionmanuscript_amd import _lib as L

    out = C.c_double()
    prec = L.PREC_FP64 if method == "fp64" else L.PREC_FP32_DC
    n = max(24, n - n % 24)  # a multiple of every depth the sweeps can run per launch
    rc = L.load().cdiff_bench_sweeps(sol._h, float(dt), prec, KERNELS[(method, sol.ndim)][2], int(n), C.byref(out))
    if rc:
        raise RuntimeError(L.load().cdiff_last_error(sol._h).decode())
    return out.value


def run_cpu(args, mesh, T0, budget_s):
    """The oracle on the first steps of the same workload, up to ~budget_s.
    Threads: OMP_NUM_THREADS (the host's CPU share per GPU where it is set),
    else every core os.cpu_count() reports."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import apt_oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    O.set_num_threads(threads)
    w = WORKLOADS[args.workload]
    s = O.OracleSolver(mesh, w["bcs"], kappa=w["kappa"])
    T = T0.copy(order="F")
    s.update_conductivity(T)
    dt = args.dt0
    t0 = time.perf_counter()
    nsteps, nit = 0, 0
    while nsteps < args.steps:
        st, dt = s.step(T, dt, rtol=0.0, atol=args.tol, max_iter=10**7)
        nsteps += 1
        nit += st["niter"]
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return dict(s_per_step=el / nsteps, steps=nsteps, iters=nit, threads=O.num_threads(), elapsed=el)


def load_traffic(kernel, n, ndim):
    """Measured HBM bytes per launch of `kernel` at n^ndim (rocprofv3 counters,
    calibrated against a copy kernel) and where they come from, if
    profiles/traffic.json has them."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
        e = d.get(f"{kernel} @ {'x'.join([str(n)] * ndim)}")
        if e:
            src = d.get("_source", {})
            return e, f"{src.get('file', 'profiles/traffic.json')}, measured {e.get('measured', src.get('measured', '?'))}"
    return None, None


def implicit_oracle_check(n=512, dt=1e-4, steps=2):
    """Error of the mixed-precision step against the FP64 implicit-direct
    oracle (oracle/implicit_oracle.py: the reference's conservative scheme,
    SciPy direct solve) on the Gaussian subsample: wavy grid A = 0.03, kappa = 1,
    zero-flux faces, dt = 1e-4.  The two are different consistent
    discretisations, so the figure is their difference over the central half of
    the domain relative to the step's increment there (O(h^2))."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import implicit_oracle as I
    from cases import gaussian_2d

    from curvilineardiffusionmanuscript_amd import solver as S

    mesh, T0 = gaussian_2d(n)
    law = S.PowerLawConductivity(1.0)
    dev = S.PseudoTransientSolver(mesh, [S.NeumannBC()] * 4)
    imp = I.ImplicitOracle2D(mesh, [("neumann", 0.0)] * 4)
    Td, Ti = T0.copy(order="F"), T0.copy(order="F")
    S.update_conductivity(dev, mesh, Td, 1.0, 1.0, law)
    q = slice(1 + n // 4, 1 + 3 * n // 4)
    errs, rels = [], []
    for _ in range(steps):
        Tprev = Td.copy()
        S.nonlinear_thermal_conduction_step(dev, mesh, Td, 1.0, 1.0, law, dt, rtol=0.0, atol=1e-8, max_iter=10**6,
                                            precision="fp32_dc", calculate_next_dt=False, show_convergence=False)
        imp.step(Ti, dt, cutoff=False)
        err = float(np.max(np.abs(Td - Ti)[q, q]))
        errs.append(err)
        rels.append(err / float(np.max(np.abs(Td - Tprev)[q, q])))
        Ti[...] = Td
    dev.close()
    return {"workload": f"2D {n}x{n} Gaussian, wavy grid A=0.03, kappa=1, zero flux, dt={dt:g}, {steps} steps",
            "max_abs_err_per_step": errs, "max_err_over_step_increment": rels,
            "region": "central half of the domain (the two schemes differ in the boundary band)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hot_disk_2d", choices=sorted(WORKLOADS),
                    help="hot_disk_2d: the headline 2D case; spots_3d: the 3D mixed-precision case")
    ap.add_argument("--n", type=int, default=None, help="cells per axis (default: 4096 in 2D, 256 in 3D)")
    ap.add_argument("--seed", type=int, default=None, help="workload seed (default: 0 in 2D, 1 in 3D)")
    ap.add_argument("--dt0", type=float, default=None, help="first time step (default: 1e-6 in 2D, 1e-5 in 3D)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--method", default="fp32_dc", choices=["fp64", "fp32_dc"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 APT comparison run")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the run with the learned warm start (2D mixed precision only)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-implicit", action="store_true",
                    help="skip the 512^2 comparison with the FP64 implicit-direct oracle")
    ap.add_argument("--sweep-samples", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: there is no multi-GPU path (the APT would shard by domain "
            "decomposition, exchanging a one-cell ring of T and the ring faces' fluxes every "
            "pseudo-iteration). Run with --gpus 1.",
            file=sys.stderr,
        )
        sys.exit(2)

    w = WORKLOADS[args.workload]
    ndim = w["ndim"]
    if args.n is None:
        args.n = w["n"]
    if args.seed is None:
        args.seed = 0 if ndim == 2 else 1
    if args.dt0 is None:
        args.dt0 = w["dt0"]
    t_setup = time.perf_counter()
    mesh, T0 = make_workload(args.workload, args.n, args.seed)
    log(f"workload {args.workload} {args.n}^{ndim} built in {time.perf_counter() - t_setup:.1f}s")
    g = run_gpu(args, mesh, T0, args.method)
    ref = None
    if args.method != "fp64" and not args.no_fp64:
        ref = run_gpu(args, mesh, T0, "fp64")
    lrn = None
    if args.method == "fp32_dc" and ndim == 2 and not args.no_learned:
        from curvilineardiffusionmanuscript_amd.learned import WarmStartNet

        net = WarmStartNet.load()
        lrn = run_gpu(args, mesh, T0, args.method, net=net)
        # for reference: the network consulted at every step from the second
        # (tuning warm_policy 1), whatever its expected saving
        lrn_all = run_gpu(args, mesh, T0, args.method, net=net, tuning={"warm_policy": 1})
        net.close()
    n_cells = args.n**ndim
    bpc = BYTES_PER_CELL[(args.method, ndim)]
    # the headline run: the mixed-precision step with the learned warm start
    # where it applies (2D), else the method itself
    head = lrn if lrn is not None else g
    s_per_step = head["elapsed"] / args.steps
    it_mean = float(np.mean(head["iters"]))
    teff = bpc * n_cells * sum(head["iters"]) / head["elapsed"] / 1e9
    s_plain = g["elapsed"] / args.steps
    log(f"GPU {args.method}: {s_plain * 1e3:.3f} ms/step, {np.mean(g['iters']):.1f} its/step, "
        f"sweep {g['sweep_ms'] * 1e3:.1f} us/iteration")
    if lrn is not None:
        log(f"GPU {args.method}+learned: {s_per_step * 1e3:.3f} ms/step, {it_mean:.1f} its/step, warm start "
            f"used on {lrn['warm_used']} of {args.steps} steps, T_eff(step) {teff:.1f} GB/s")
    extra = {}
    if lrn is not None:
        extra["warm_start_used_steps"] = lrn["warm_used"]
        # the network is consulted only where the saving expected from it
        # exceeds what a consultation costs (on the mixed-precision step it
        # does not: the steps are then the plain ones)
        extra["warm_start_consulted_steps"] = lrn["warm_consulted"]
        extra["learned_consulted_every_step"] = {
            "s_per_step": round(lrn_all["elapsed"] / args.steps, 6),
            "iters_per_step": round(float(np.mean(lrn_all["iters"])), 2),
            "outer_per_step": round(float(np.mean(lrn_all["outer"])), 2),
            "warm_start_used_steps": lrn_all["warm_used"],
            "consulted_steps": lrn_all["warm_consulted"],
        }
        extra["without_learned_warm_start"] = {
            "s_per_step": round(s_plain, 6),
            "iters_per_step": round(float(np.mean(g["iters"])), 2),
            "outer_per_step": round(float(np.mean(g["outer"])), 2),
            "max_abs_err": max(g["abs_err"]),
        }
    if ref is not None:
        s64 = ref["elapsed"] / args.steps
        scale = float(np.max(np.abs(ref["T_end"])))
        err = float(np.max(np.abs(head["T_end"] - ref["T_end"]))) / scale
        log(f"GPU fp64: {s64 * 1e3:.3f} ms/step, {np.mean(ref['iters']):.1f} its/step; speedup "
            f"{s64 / s_per_step:.2f}x{f' (without the network {s64 / s_plain:.2f}x)' if lrn is not None else ''}; "
            f"max|T - T_fp64|/max|T| = {err:.2e}")
        extra.update({
            "fp64_s_per_step": round(s64, 6),
            "fp64_iters_per_step": round(float(np.mean(ref["iters"])), 2),
            "speedup_vs_fp64": round(s64 / s_per_step, 3),
            "max_rel_diff_vs_fp64": err,
        })
        if lrn is not None:
            w = extra["without_learned_warm_start"]
            w["speedup_vs_fp64"] = round(s64 / s_plain, 3)
            w["max_rel_diff_vs_fp64"] = float(np.max(np.abs(g["T_end"] - ref["T_end"]))) / scale

    kname, per_launch, _ = KERNELS[(args.method, ndim)]
    per_it_ms = g["sweep_ms"]
    launch_ms = per_it_ms * per_launch if per_it_ms == per_it_ms else None
    comp_bytes = COMPULSORY_BYTES_PER_CELL[(args.method, ndim)] * n_cells
    aeff_bytes = bpc * n_cells * per_launch
    achieved = comp_bytes / (launch_ms / 1e3) / 1e9 if launch_ms else None
    t_eff = aeff_bytes / (launch_ms / 1e3) / 1e9 if launch_ms else None
    traffic, traffic_src = load_traffic(kname, args.n, ndim)
    roofline = {
        "kernel": kname,
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "traffic_source": traffic_src,
        "compulsory_bytes_per_launch": comp_bytes,
        "compulsory_bytes_per_cell": COMPULSORY_BYTES_PER_CELL[(args.method, ndim)],
        "iterations_per_launch": per_launch,
        "mean_launch_us": round(launch_ms * 1e3, 2) if launch_ms else None,
        "t_eff": {"convention": f"A_eff = {bpc} B per cell and pseudo-iteration x {per_launch} iterations per launch",
                  "GBs": round(t_eff, 1) if t_eff else None,
                  "frac": round(t_eff / HBM_PEAK_GBS, 4) if t_eff else None},
    }
    if traffic and launch_ms:
        roofline["traffic_GBs"] = round(traffic["bytes_per_launch"] / (launch_ms / 1e3) / 1e9, 1)
    cpu = None
    if not args.no_cpu:
        c = run_cpu(args, mesh, T0, args.cpu_seconds)
        k = c["steps"]
        same = {"gpu_" + name + "_s_per_step": round(float(np.mean(r["step_times"][:k])), 6)
                for name, r in (("head", head), ("fp64", ref)) if r is not None}
        gpu_it = (float(np.sum(ref["step_times"][:k])) / float(np.sum(ref["iters"][:k])) * 1e3) if ref else None
        log(f"CPU oracle: {c['s_per_step']:.3f} s/step over {k} steps ({c['iters']} iterations), {c['threads']} "
            f"threads; the same steps on the GPU: {same}")
        cpu = {
            "value": round(c["s_per_step"], 5),
            "unit": "s/time-step",
            "cores": c["threads"],
            "kind": "port",
            "sample": f"first {k} of the {args.steps} steps ({c['iters']} pseudo-iterations), same inputs and "
            f"tolerance, FP64, oracle/apt_oracle.c with OpenMP on OMP_NUM_THREADS threads",
            "same_steps": same,
            "ms_per_iteration": round(c["elapsed"] / c["iters"] * 1e3, 4),
            "gpu_fp64_ms_per_iteration": round(gpu_it, 4) if gpu_it else None,
        }
        if ndim == 2 and not args.no_implicit:
            t_i = time.perf_counter()
            extra["error_vs_implicit_oracle"] = implicit_oracle_check()
            log(f"implicit-direct oracle check: {extra['error_vs_implicit_oracle']['max_err_over_step_increment']} "
                f"({time.perf_counter() - t_i:.1f} s)")
    out = {
        "metric": "s/time-step + T_eff %HBM peak: FP32+learned vs FP64 APT, same 1e-8 residual",
        "value": round(s_per_step, 6),
        "unit": "s/time-step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_step * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.method == "fp64" else "f32+f64",
        "data": "synthetic (seeded hot disk, analytic wavy mesh)" if ndim == 2
        else "synthetic (seeded Gaussian spots, analytic wavy mesh)",
        "config": {
            "workload": (
                f"2D {args.n}x{args.n} wavy grid (A=0.05), kappa=T^2.5, hot disk seed {args.seed}, "
                f"Dirichlet T=1 left, zero-flux elsewhere, adaptive dt0={args.dt0:g}, atol={args.tol:g}, "
                f"method={args.method}{'+learned_warm_start' if lrn is not None else ''}"
                if ndim == 2
                else f"3D {args.n}^3 wavy grid (A=0.04), kappa=T^3, 8 Gaussian spots seed {args.seed}, "
                f"zero-flux, adaptive dt0={args.dt0:g}, atol={args.tol:g}, method={args.method}"
            ),
            "iters_per_step": round(it_mean, 2),
            "outer_per_step": round(float(np.mean(head["outer"])), 2),
            "max_abs_err": max(head["abs_err"]),
            "t_eff_step_GBs": round(teff, 1),
            "t_eff_step_frac": round(teff / HBM_PEAK_GBS, 4),
            **extra,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
