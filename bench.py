"""Benchmark of the fused esoteric-pull step on one MI355X.

Workload: D3Q27 factorised central moments, FP32 shifted populations, 512^3
periodic Taylor-Green vortex (u0 = 0.04, Re = 1600 on L = D/2pi, seeded uniform
perturbation 1e-4), the configuration the throughput metric is quoted on.  A
step is one stream+collide update of every node; the library runs the steps of
this box in pairs, one pass over the population array per pair (a persistent
launch and a short launch for the rows it defers), and the rate is steps times
nodes over the time of the whole sequence.

Prints ONE JSON line on stdout: MLUPS over interior nodes (the reference's
convention, src/app/main.cpp:50), the HBM roofline of the step kernel measured
with HIP events on the solver's stream (216 B per node-step FP32, read +
write), and the CPU restatement timed on this host's cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, published
FULL_SERVING_REQUESTS = 4000  # the serving workload's stated request count
DTYPES = {"f64": lambda abi: abi.F64, "f32": lambda abi: abi.F32, "f16": lambda abi: abi.F16,
          "bf16": lambda abi: abi.BF16, "i16": lambda abi: abi.I16,
          "e2m13": lambda abi: abi.E2M13}
FUSION_MODES = {"auto": lambda abi: abi.FUSION_AUTO, "off": lambda abi: abi.FUSION_OFF,
                "on": lambda abi: abi.FUSION_ON}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def taylor_green_config(abi, D, dtype, op, storage=None):
    u0, Re = 0.04, 1600.0
    nu = u0 * (D / (2.0 * math.pi)) / Re
    omega = 1.0 / (3.0 * nu + 0.5)
    storage = abi.SHIFTED if storage is None else storage
    return abi.make_config(abi.D3Q27, op, storage, dtype, (D, D, D), omega=omega), u0


def bytes_per_node(abi, dtype, Q=27):
    return 2 * Q * {abi.F64: 8, abi.F32: 4, abi.F16: 2, abi.BF16: 2, abi.I16: 2,
                    abi.E2M13: 2}[dtype]


def gpu_leg(D, steps, warmup, dtype, op, seed=3, field=None, fusion=None):
    from m3lb_amd import abi, cases
    from m3lb_amd.solver import FluidSolver
    cfg, u0 = taylor_green_config(abi, D, dtype, op)
    t0 = time.time()
    s = FluidSolver(cfg)
    if fusion is not None:
        s.set_step_fusion(fusion)
    if field is None:
        field = cases.tgv3d_benchmark(cfg, u0, perturbation=1e-4, seed=seed)
    s.initialize_field(*field)
    log(f"[bench] init {D}^3 {abi.DTYPE_NAMES[dtype]} op={abi.OP_NAMES[op]} in {time.time() - t0:.1f}s")
    m0 = s.total_mass()
    s.time_steps(warmup)
    ms = s.time_steps(steps)
    m1 = s.total_mass()
    e, w, finite = s.energy_enstrophy()
    s.close()
    nodes = D ** 3
    sec = ms / 1e3
    mlups = nodes * steps / sec / 1e6
    per_launch_s = sec / steps
    achieved = bytes_per_node(abi, dtype) * nodes / per_launch_s / 1e9
    return dict(mlups=mlups, ms=ms, ms_per_step=ms / steps, achieved_gbs=achieved,
                mass_drift=abs(m1 - m0) / m0, finite=finite)


def host_cpu():
    """Cores this process may run on, the host's logical CPU count and model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, os.cpu_count() or 1, model


def cpu_threads():
    """Every core the job is given: the affinity set, bounded by the job's
    OMP_NUM_THREADS share when the host sets one."""
    usable, _, _ = host_cpu()
    share = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(usable, int(share))) if share.isdigit() and int(share) > 0 else usable


def cpu_leg(D, steps, threads, field=None):
    """The oracle (CPU restatement of the reference step) on this host, timed
    as src/app/main.cpp:41-51: one warm-up step, then `steps` steps."""
    from m3lb_amd import abi, cases
    from oracle.binding import OracleSolver
    cfg, u0 = taylor_green_config(abi, D, abi.F32, abi.OP_CM)
    o = OracleSolver(cfg, threads=threads)
    if field is None:
        field = cases.tgv3d_benchmark(cfg, u0, perturbation=1e-4, seed=3)
    o.initialize_field(*field)
    sec = o.time_steps(steps)
    o.close()
    return D ** 3 * steps / sec / 1e6, sec


def _pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q)) if v else None


def serving_batched(work, budget_bytes, tune_steps, steps_per_chunk, max_resident_nodes=0,
                    use_graph=True, dtype=None):
    """The requests through one batched server: arrivals replayed on the
    server clock, continuous admission, completion times from the server."""
    from m3lb_amd import abi
    from m3lb_amd.serve import Server
    scfg = abi.make_server_config(budget_bytes=budget_bytes, max_boxes=4096,
                                  steps_per_chunk=steps_per_chunk,
                                  max_resident_nodes=max_resident_nodes, use_graph=use_graph,
                                  dtype=abi.F32 if dtype is None else dtype)
    n = len(work)
    with Server(scfg) as srv:
        tune = srv.autotune(steps=tune_steps) if tune_steps > 0 else None
        log(f"[bench] serving autotune: {tune}")
        t0 = srv.clock() + 0.01
        nxt, ttc, last, peak_res = 0, [], t0, 0
        next_log = t0 + 20.0
        while len(ttc) < n:
            now = srv.clock()
            if now > next_log:
                log(f"[bench] serving batched: {len(ttc)}/{n} complete at {now - t0:.0f} s")
                next_log = now + 20.0
            while nxt < n and t0 + work[nxt][0] <= now:
                req = work[nxt][1]
                req.arrival = t0 + work[nxt][0]
                req.keep = 0
                srv.submit(req)
                nxt += 1
            if nxt < n:
                wait = int(min(0.005, max(0.0, t0 + work[nxt][0] - srv.clock())) * 1e6)
                srv.progress(wait_us=max(wait, 1))
            else:
                srv.progress(wait_us=-1)
            for r in srv.poll():
                ttc.append(r["completed"] - r["arrival"])
                last = max(last, r["completed"])
            peak_res = max(peak_res, srv.stats()["resident"])
        st = srv.stats()
    lu = sum(r.nx * r.ny * r.nz * r.steps for _, r in work)
    span = last - t0
    return {"mlups": lu / span / 1e6, "makespan_s": span, "p50_s": _pct(ttc, 50),
            "p99_s": _pct(ttc, 99), "device_mlups": st["lattice_updates"] / st["device_seconds"] / 1e6,
            "peak_resident": peak_res, "peak_bytes": st["bytes_peak"], "chunks": st["chunks"],
            "steps_per_chunk": steps_per_chunk, "max_resident_nodes": max_resident_nodes,
            "tune": tune, "lattice_updates": lu}


def serving_per_case(work, dtype=None):
    """The same requests one case at a time, in arrival order, each on its own
    single-box solver: create, geometry, initialise, one launch per step,
    read the mass (synchronises), destroy."""
    from m3lb_amd import abi, cases
    from m3lb_amd.solver import FluidSolver
    scfg = abi.make_server_config(dtype=abi.F32 if dtype is None else dtype)
    t0 = time.perf_counter() + 0.01
    free, ttc, step_s = t0, [], 0.0
    next_log = t0 + 20.0
    for off, r in work:
        if time.perf_counter() > next_log:
            log(f"[bench] serving per case: {len(ttc)}/{len(work)} complete")
            next_log = time.perf_counter() + 20.0
        arr = t0 + off
        now = time.perf_counter()
        if now < arr:
            time.sleep(arr - now)
        with FluidSolver(abi.request_config(scfg, r)) as s:
            if r.walls == abi.WALLS_Y:
                s.set_geometry(cases.channel_flags((r.nx, r.ny, r.nz)))
            if r.init == abi.INIT_TAYLOR_GREEN:
                s.initialize_taylor_green(r.u[0])
            else:
                s.initialize(r.rho0, *r.u)
            step_s += s.time_steps(r.steps) / 1e3
            s.total_mass()
        free = time.perf_counter()
        ttc.append(free - arr)
    lu = sum(r.nx * r.ny * r.nz * r.steps for _, r in work)
    span = free - t0
    return {"mlups": lu / span / 1e6, "makespan_s": span, "p50_s": _pct(ttc, 50),
            "p99_s": _pct(ttc, 99), "step_only_mlups": lu / step_s / 1e6}


PRECISION_FORMS = ("f64", "f32", "f16", "bf16", "i16", "e2m13")


def precision_study(D, steps, seed=3, forms=PRECISION_FORMS):
    """The Taylor-Green workload stored in FP64, FP32 and every 16-bit form from the
    same initial field: velocity relative L2 and kinetic-energy difference of
    each against the FP64-storage run after `steps` steps, and MLUPS."""
    import numpy as np
    from m3lb_amd import abi, cases
    from m3lb_amd.solver import FluidSolver
    cfg0, u0 = taylor_green_config(abi, D, abi.F64, abi.OP_CM)
    field = cases.tgv3d_benchmark(cfg0, u0, perturbation=1e-4, seed=seed)
    ref, out = None, {}
    dts = {"f64": abi.F64, "f32": abi.F32, "f16": abi.F16, "bf16": abi.BF16, "i16": abi.I16,
           "e2m13": abi.E2M13}
    for name in forms:
        dt = dts[name]
        cfg, _ = taylor_green_config(abi, D, dt, abi.OP_CM)
        with FluidSolver(cfg) as s:
            s.initialize_field(*field)
            e0, _, _ = s.energy_enstrophy()
            ms = s.time_steps(steps)
            _, ux, uy, uz = s.compute_macroscopic()
            e1, w1, fin = s.energy_enstrophy()
            scales = s.scales() if dt not in (abi.F64, abi.F32) else None
        row = {"mlups": round(D ** 3 * steps / (ms / 1e3) / 1e6, 1), "energy": e1,
               "energy_ratio_to_initial": e1 / e0, "enstrophy": w1, "finite": bool(fin)}
        if ref is None:
            ref = (ux, uy, uz, e1)
        else:
            num = float(np.sum((ux - ref[0]) ** 2 + (uy - ref[1]) ** 2 + (uz - ref[2]) ** 2))
            den = float(np.sum(ref[0] ** 2 + ref[1] ** 2 + ref[2] ** 2))
            row["vel_rel_l2_vs_f64"] = (num / den) ** 0.5
            row["energy_rel_diff_vs_f64"] = abs(e1 - ref[3]) / ref[3]
        if scales is not None:
            row["scale_log2"] = [int(round(math.log2(v))) for v in scales]
        out[name] = row
        log(f"[bench] precision {name}: {row}")
        del ux, uy, uz
    return {"size": D, "steps": steps, "seed": seed, "rows": out}


def configs_leg():
    """The other synthetic workloads of the benchmark definition, one line
    each (not part of the headline):
      C0  D3Q27 BGK (second-order equilibrium) FP64 64^3 Taylor-Green
          (benchmark form) u0 = 0.05 at
          Re 1600 on L/(2 pi), seeded U(-1e-4, 1e-4) perturbation (seed 0),
          500 steps, E and Omega every 50;
      C1  D3Q27 CM FP32 256^3, u0 = 0.04, Re 1600, seed 1, 1000 steps, mass
          drift and E every 100;
      C2  D3Q27 CM FP32 512 x 256 x 256 channel: Solid y and z layers, 200
          spheres r ~ U[6, 14] (seed 2), Guo g_x = 1e-6, tau = 7/8 (the CM
          magic value), from rest, 5000 steps; mean u_x = sum rho u_x / sum
          rho and the momentum-exchange drag on the spheres."""
    import numpy as np
    from m3lb_amd import abi, cases
    from m3lb_amd.solver import FluidSolver
    out = {}
    # C0
    L, u0 = 64, 0.05
    nu = u0 * L / (2.0 * math.pi * 1600.0)
    cfg = abi.make_config(abi.D3Q27, abi.OP_BGK, abi.SHIFTED, abi.F64, (L, L, L),
                          omega=1.0 / (3.0 * nu + 0.5))
    with FluidSolver(cfg) as s:
        s.initialize_field(*cases.tgv3d_benchmark(cfg, u0, perturbation=1e-4, seed=0))
        e0, w0, _ = s.energy_enstrophy()
        rows, ms = [], 0.0
        diverged = None
        for k in range(10):
            ms += s.time_steps(50)
            e, w, fin = s.energy_enstrophy()
            rows.append([50 * (k + 1), e / e0, w / w0])
            if not fin:
                diverged = 50 * (k + 1)
                break
        out["C0"] = {"steps": rows[-1][0], "mlups": round(L ** 3 * rows[-1][0] / ms / 1e3, 1),
                     "E_over_E0_and_Omega_over_Omega0": rows, "diverged_at": diverged,
                     "tau": 3.0 * nu + 0.5}
    log(f"[bench] C0: {out['C0']['mlups']} MLUPS, final E/E0 {rows[-1][1]:.6f}, "
        f"diverged {diverged}")
    # C1
    L, u0 = 256, 0.04
    nu = u0 * L / (2.0 * math.pi * 1600.0)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (L, L, L),
                          omega=1.0 / (3.0 * nu + 0.5))
    with FluidSolver(cfg) as s:
        s.initialize_field(*cases.tgv3d_benchmark(cfg, u0, perturbation=1e-4, seed=1))
        m0 = s.total_mass()
        e0, _, _ = s.energy_enstrophy()
        rows, ms = [], 0.0
        for k in range(10):
            ms += s.time_steps(100)
            e, _, fin = s.energy_enstrophy()
            rows.append([100 * (k + 1), (s.total_mass() - m0) / m0, e / e0])
        out["C1"] = {"steps": 1000, "mlups": round(L ** 3 * 1000 / ms / 1e3, 1),
                     "mass_drift_and_E_over_E0": rows, "finite": bool(fin),
                     "tau": 3.0 * nu + 0.5}
    log(f"[bench] C1: {out['C1']['mlups']} MLUPS, mass drift {rows[-1][1]:.2e}")
    # C2
    shape = (512, 256, 256)
    t0 = time.time()
    flags, body, spheres = cases.sphere_channel(shape, n_spheres=200, rmin=6.0, rmax=14.0, seed=2)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, shape,
                          periodic=(True, False, False), omega=1.0 / 0.875,
                          force=(1e-6, 0.0, 0.0))
    with FluidSolver(cfg) as s:
        s.set_geometry(flags)
        s.initialize(1.0)
        log(f"[bench] C2 geometry: {len(spheres)} spheres, solid fraction "
            f"{float(np.mean(flags == abi.SOLID)):.4f}, set up in {time.time() - t0:.1f}s")
        ms = s.time_steps(5000)
        mass = s.total_mass()
        mom = s.total_momentum(0)
        drag = s.momentum_exchange(body)
        fluid = int(np.sum(flags == abi.FLUID))
        out["C2"] = {"steps": 5000, "mlups": round(np.prod(shape) * 5000 / ms / 1e3, 1),
                     "spheres": len(spheres), "solid_fraction": float(np.mean(flags == abi.SOLID)),
                     "mean_ux": mom / mass, "drag_on_spheres": list(drag),
                     "body_force_on_fluid": 1e-6 * fluid, "tau": 0.875}
    log(f"[bench] C2: {out['C2']['mlups']} MLUPS, mean u_x {mom / mass:.4e}, drag x {drag[0]:.4e}")
    return out


def serving_leg(n, rate, seed, budget_gb, tune_steps, steps_per_chunk, cap=0, use_graph=True,
                edges=(32, 192), dtype_name="f32"):
    from m3lb_amd import abi, cases
    dtype = DTYPES[dtype_name](abi)
    work = cases.serving_workload(n, seed=seed, rate=rate, edges=edges)
    b = serving_batched(work, int(budget_gb * 1e9), tune_steps, steps_per_chunk, cap,
                        use_graph, dtype)
    log(f"[bench] serving batched: {b['mlups']:.0f} MLUPS, p50 {b['p50_s']:.2f} s, "
        f"p99 {b['p99_s']:.2f} s, makespan {b['makespan_s']:.1f} s, peak resident "
        f"{b['peak_resident']}")
    work = cases.serving_workload(n, seed=seed, rate=rate, edges=edges)
    c = serving_per_case(work, dtype)
    log(f"[bench] serving one launch per case: {c['mlups']:.0f} MLUPS, p50 {c['p50_s']:.2f} s, "
        f"p99 {c['p99_s']:.2f} s, makespan {c['makespan_s']:.1f} s")
    rnd = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
    sample = (f"the first {n} requests of the seeded arrival stream; the full "
              f"{FULL_SERVING_REQUESTS}-request workload is not measured by this run"
              if n < FULL_SERVING_REQUESTS and edges == (32, 192) else f"{n} requests")
    return {"requests": n, "arrival_rate_per_s": rate, "seed": seed,
            "workload": f"D3Q27 CM {dtype_name.upper()} shifted; 50% periodic Taylor-Green u0=0.04, 50% "
                        "bounce-back channels g_x~U[1e-7,1e-5]; edges log-uniform "
                        f"[{edges[0]},{edges[1]}) "
                        f"per axis; tau~U[0.51,0.8]; steps~U{{200..3000}}; Poisson arrivals; "
                        + sample,
            "budget_bytes": int(budget_gb * 1e9), "batched": rnd(b), "per_case": rnd(c),
            "speedup_mlups": round(b["mlups"] / c["mlups"], 3),
            # share of the one-case-at-a-time run's wall time its step launches
            # kept the device busy: with the same step rate, batching can at
            # most remove the rest
            "per_case_busy_fraction": round(c["mlups"] / c["step_only_mlups"], 4),
            "p99_ratio_batched_to_per_case": round(b["p99_s"] / c["p99_s"], 3)}


def serving_small(args):
    """The serving comparison on small boxes (edges [16, 48)), where one launch
    per step per case is launch-bound."""
    if args.serving_small_requests <= 0 or args.serving_requests <= 0:
        return None
    return serving_leg(args.serving_small_requests, args.serving_small_rate, 5,
                       args.serving_budget_gb, args.serving_tune_steps, args.serving_chunk,
                       args.serving_cap, not args.serving_no_graph, edges=(16, 48))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "i16", "e2m13", "f64"])
    ap.add_argument("--cpu-size", type=int, default=0,
                    help="box edge of the CPU baseline (default: the GPU workload's)")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="timed CPU steps after one warm-up (a bounded sample)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fusion", default="auto", choices=sorted(FUSION_MODES),
                    help="two steps per pass: the library's choice (auto), never (off: one "
                         "launch per step) or whenever the box allows (on)")
    ap.add_argument("--extra", action="store_true",
                    help="also time the 16-bit and FP64 storage forms and BGK (stderr)")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests of the batched-serving leg (0 skips it)")
    ap.add_argument("--serving-rate", type=float, default=40.0, help="arrivals per second")
    ap.add_argument("--serving-budget-gb", type=float, default=200.0)
    ap.add_argument("--serving-chunk", type=int, default=32,
                    help="steps per chunk (32: best of 4, 8, 16 and 32 on both serving samples)")
    ap.add_argument("--serving-tune-steps", type=int, default=3)
    ap.add_argument("--serving-cap", type=int, default=2400000,
                    help="max resident interior nodes of the server (0: memory only); "
                         "2.4 M D3Q27 FP32 nodes keep the resident populations in the "
                         "256 MB last-level cache")
    ap.add_argument("--serving-small-requests", type=int, default=400,
                    help="requests of the small-box serving workload (edges [16, 48)); 0 skips it")
    ap.add_argument("--serving-small-rate", type=float, default=400.0, help="arrivals per second")
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--serving-dtype", default="f32", choices=sorted(DTYPES),
                    help="storage of the served boxes (--serving-only; 16-bit forms shifted)")
    ap.add_argument("--no-e2m13", action="store_true",
                    help="skip the E2M13-storage run of the same workload")
    ap.add_argument("--e2m13-steps", type=int, default=2000,
                    help="steps of the E2M13 and FP64-storage runs compared in the bench line")
    ap.add_argument("--serving-no-graph", action="store_true",
                    help="launch the server's chunks directly instead of replaying a "
                         "captured hipGraph (for kernel-trace profiles)")
    ap.add_argument("--configs", action="store_true",
                    help="only run the C0/C1/C2 workloads (Taylor-Green FP64 BGK 64^3, CM "
                         "FP32 256^3, sphere channel 512x256x256) and print their JSON")
    ap.add_argument("--precision-study", type=int, default=0, metavar="STEPS",
                    help="only run the FP64/FP32/16-bit storage study for STEPS steps")
    ap.add_argument("--stream-probe-only", type=int, default=0, metavar="BYTES",
                    help="only run the plain in-place stream of BYTES (counter calibration)")
    args = ap.parse_args()

    if args.stream_probe_only:
        from m3lb_amd import solver
        ms = solver.stream_probe(args.stream_probe_only, nontemporal=False, reps=3)
        log(f"[bench] stream probe {args.stream_probe_only} B in place: {ms:.3f} ms, "
            f"{2 * args.stream_probe_only / ms / 1e6:.0f} GB/s")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the step shards by z-slabs across GPUs by design, not implemented here)",
              file=sys.stderr)
        sys.exit(2)

    from m3lb_amd import abi
    if args.precision_study:
        print(json.dumps({"precision": precision_study(args.size, args.precision_study)}),
              flush=True)
        return
    if args.configs:
        print(json.dumps({"configs": configs_leg()}), flush=True)
        return
    if args.serving_only:
        sv = serving_leg(args.serving_requests, args.serving_rate, 4, args.serving_budget_gb,
                         args.serving_tune_steps, args.serving_chunk, args.serving_cap,
                         not args.serving_no_graph, dtype_name=args.serving_dtype)
        small = serving_small(args)
        print(json.dumps({"serving": sv, "serving_small_boxes": small}), flush=True)
        return
    dt = {"f32": abi.F32, "f16": abi.F16, "bf16": abi.BF16, "i16": abi.I16,
          "e2m13": abi.E2M13, "f64": abi.F64}[args.dtype]
    from m3lb_amd import cases
    cfg0, u00 = taylor_green_config(abi, args.size, abi.F32, abi.OP_CM)
    field = cases.tgv3d_benchmark(cfg0, u00, perturbation=1e-4, seed=3)
    r = gpu_leg(args.size, args.steps, args.warmup, dt, abi.OP_CM, field=field,
                fusion=FUSION_MODES[args.fusion](abi))
    log(f"[bench] gpu {args.size}^3 CM {args.dtype}: {r['mlups']:.0f} MLUPS, "
        f"{r['ms_per_step']:.3f} ms/step, {r['achieved_gbs']:.0f} GB/s, "
        f"mass drift {r['mass_drift']:.2e}, finite {r['finite']}")

    if args.extra:
        for name, d2, op in (("f16", abi.F16, abi.OP_CM), ("bf16", abi.BF16, abi.OP_CM),
                             ("i16", abi.I16, abi.OP_CM), ("e2m13", abi.E2M13, abi.OP_CM),
                             ("f64", abi.F64, abi.OP_CM), ("f32-bgk", abi.F32, abi.OP_BGK)):
            x = gpu_leg(args.size, max(10, args.steps // 2), 5, d2, op)
            log(f"[bench] extra {name}: {x['mlups']:.0f} MLUPS {x['achieved_gbs']:.0f} GB/s "
                f"({x['achieved_gbs'] / PEAK_HBM_GBS:.3f} of peak)")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_threads()
        usable, logical, model = host_cpu()
        csize = args.cpu_size or args.size
        mlups_cpu, sec = cpu_leg(csize, args.cpu_steps, threads,
                                 field=field if csize == args.size else None)
        cpu = {"value": round(mlups_cpu, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
               "host_logical_cpus": logical, "job_cpus": usable, "cpu_model": model,
               "sample": f"oracle/ C restatement of the reference step (Kokkos Threads "
                         f"unavailable offline), D3Q27 CM FP32 shifted, the same Taylor-Green "
                         f"workload at {csize}^3, {args.cpu_steps} timed steps after 1 warm-up "
                         f"({sec:.1f} s), {threads} threads (every core given to this job; "
                         f"the host has {logical} logical CPUs, {model})"}
        log(f"[bench] cpu {csize}^3: {mlups_cpu:.2f} MLUPS on {threads} threads "
            f"({logical} logical CPUs on the host, {model})")
    del field

    bpn = bytes_per_node(abi, dt)
    # The device's streaming ceiling, measured now on an array the size of the
    # population array: the best a no-reuse in-place kernel reaches here.
    from m3lb_amd import solver
    arr = bpn // 2 * args.size ** 3
    ceil = {}
    for nt in (False, True):
        ms = solver.stream_probe(arr, nontemporal=nt, reps=3)
        ceil["nontemporal" if nt else "plain"] = round(2 * arr / ms / 1e6, 1)
    best_ceiling = max(ceil.values())
    log(f"[bench] in-place stream of {arr / 1e9:.2f} GB: plain {ceil['plain']:.0f} GB/s, "
        f"nontemporal {ceil['nontemporal']:.0f} GB/s; step kernel at "
        f"{r['achieved_gbs'] / best_ceiling:.3f} of the better")

    # Bytes per step between the L2 caches and the memory fabric, from the
    # committed rocprofv3 --pmc summary of this exact workload (FETCH_SIZE +
    # WRITE_SIZE of one two-step pass, calibrated on a copy kernel, halved),
    # used only while the loaded library holds the kernels it was measured on.
    # These counters see requests that the memory-side 256 MB cache serves as
    # well as those that reach HBM.
    traffic, traffic_src = None, None
    tj = os.path.join(ROOT, "profiles", "k_pair_cm_f32_512cubed_traffic.json")
    build = solver.kernel_build_id()
    if args.size == 512 and args.dtype == "f32" and args.fusion != "off" and os.path.exists(tj):
        with open(tj) as fh:
            tr = json.load(fh)
        if tr.get("build_id") == build:
            traffic = tr["traffic_bytes"]
            traffic_src = f"{tr['source']} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; build {build})"
        else:
            traffic_src = (f"null: the stored count was taken on build {tr.get('build_id')}, "
                           f"the loaded library is build {build}")

    # the same workload with E2M13 storage (108 B per node-step), run for
    # 2000 steps beside the same run with FP64 storage: rate and error
    narrow = None
    if args.dtype == "f32" and not args.no_e2m13:
        st = precision_study(args.size, args.e2m13_steps, forms=("f64", "e2m13"))
        x = st["rows"]["e2m13"]
        narrow = {"mlups": x["mlups"], "vs_f32": round(x["mlups"] / r["mlups"], 3),
                  "achieved_gbs": round(x["mlups"] * 108 / 1e3, 1), "bytes_per_node_step": 108,
                  "steps": args.e2m13_steps, "vel_rel_l2_vs_f64": x["vel_rel_l2_vs_f64"],
                  "energy_rel_diff_vs_f64": x["energy_rel_diff_vs_f64"], "finite": x["finite"]}
        log(f"[bench] e2m13 storage: {x['mlups']:.0f} MLUPS ({narrow['vs_f32']}x FP32), velocity "
            f"rel-L2 vs FP64 {x['vel_rel_l2_vs_f64']:.2e} after {args.e2m13_steps} steps")

    serving = None
    if args.serving_requests > 0:
        serving = serving_leg(args.serving_requests, args.serving_rate, 4,
                              args.serving_budget_gb, args.serving_tune_steps, args.serving_chunk,
                              args.serving_cap, not args.serving_no_graph)
    serving_small_boxes = serving_small(args)

    out = {
        "metric": "D3Q27 CM MLUPS + %HBM roofline at 512^3; batched-solve p50/p99 latency",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f16": "f16-storage/f32", "bf16": "bf16-storage/f32",
                  "i16": "i16-storage/f32", "e2m13": "e2m13-storage/f32", "f64": "f64"}[args.dtype],
        "data": "synthetic",
        "config": {"workload": f"D3Q27 central-moment esoteric-pull step, {args.dtype} shifted "
                               f"populations, {args.size}^3 periodic Taylor-Green u0=0.04 "
                               f"Re=1600 + seeded 1e-4 perturbation",
                   "nodes": args.size ** 3, "bytes_per_node_step": bpn,
                   "step_fusion": args.fusion},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(r["achieved_gbs"] / PEAK_HBM_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "traffic_source": traffic_src,
        "kernel_build_id": build,
        "stream_ceiling_gbs": ceil,
        "frac_of_stream_ceiling": round(r["achieved_gbs"] / best_ceiling, 4),
        "mass_drift": r["mass_drift"],
        "storage_e2m13": narrow,
        "serving": serving,
        "serving_small_boxes": serving_small_boxes,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
