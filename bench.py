#!/usr/bin/env python3
"""Benchmark: MLUPS of the FP64 two-phase time step on a 256^3 synthetic porous medium, one MI355X.

Workload (the headline configuration): 256^3 lattice, randomly placed overlapping solid spheres of radius
uniform in [6, 12] (SplitMix64 seed 2024) added until the porosity of the 254x254x240 sample reaches 0.22,
8 fluid buffer layers at each z end, solid walls at x and y ends, D3Q19 MRT colour-gradient model with two
population sets, velocity inlet (mean 1e-4, rectangular-duct profile) and convective outlet, contact angle
150 degrees measured through the invading non-wetting fluid 1, i.e. theta = 30 in the reference's convention
(the angle through fluid 2, which its control file requires below 90 degrees; theta = 150 there makes this
drainage diverge near step 2250), viscosity ratio 1 (nu = 0.05), surface tension 0.05 (Ca = 1e-4), drainage
start.

A "step" is one full time step: collision/streaming of both colours, boundary conditions and the
colour-gradient chain, exactly as the reference's main_iteration_kernel.  MLUPS counts fluid nodes
(pore_sum * steps / t / 1e6, the reference's formula).  Inputs are resident in HBM before timing starts.

roofline: the dominant kernel (collision + streaming) timed with HIP events on its own stream over the timed
steps; achieved = 625 B per fluid node (the algorithmic bytes of one two-phase FP64 step) * fluid nodes /
mean kernel duration.  cpu_baseline: the CPU oracle (C++/OpenMP restatement of the same step) on the host
cores for a bounded number of steps of the same workload.

``--workload singlephase`` measures the single-phase program's step instead (D3Q19 MRT, one population set,
bounce-back-optimised rates, nu = 0.1) on the same 256^3 pack, z-periodic and driven by a body force of 1e-6
(a permeability run); 305 B per fluid node.  ``--precision fp32`` stores the populations in FP32 (single phase:
deviations from the rest equilibrium, 153 B per fluid node; two phase: f + g - REST and f - g, 321 B) and
collides in FP32; ``--precision mixed`` keeps the FP32 storage and collides in FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mf-lbm_amd"))

# the headline metric as the benchmark definition names it; MLUPS of the two-phase step
METRIC = "MLUPS (D3Q19+D3Q7, FP64) at 256^3 porous; time-to-converged permeability"
B_ALG = 625.0          # bytes per fluid node per step, two-phase FP64
B_ALG32 = 321.0        # two-phase FP32 storage: 2 x 19 x 4 B x (read + write) + phi 8 + 8 + wall mask 1
B_ALG_SP = 305.0       # bytes per fluid node per step, single-phase FP64
B_ALG_SP32 = 153.0     # bytes per fluid node per step, single-phase FP32 deviation storage
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
DEVELOP_STEPS = 0      # two-phase: untimed steps before the warm-up


def workload(n: int = 256, precision: str = "fp64"):
    from mflbm import Case, geometry
    walls = geometry.sphere_pack(n, n, n, 0.22, 6.0, 12.0, 2024, buffer=8)
    ctl = {"lattice_dimensions": (n, n, n), "periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 0), "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 0.05,
           "fluid2_viscosity": 0.05, "surface_tension": 0.05, "theta": 30.0, "capillary_number": 1e-4,
           "initial_fluid_distribution_option": 1, "initial_interface_position": 8.0,
           "saturation_injection": 1.0, "excluded_layers": (8, 8)}
    return Case(ctl, walls, precision=precision)


def workload_sp(n: int = 256, precision: str = "fp64"):
    from mflbm import geometry
    from mflbm.singlephase import SPCase
    walls = geometry.sphere_pack(n, n, n, 0.22, 6.0, 12.0, 2024, buffer=8)
    ctl = {"lattice_dimensions": (n, n, n), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-6,
           "MRT_collision_parameter_preset": 1}
    return SPCase(ctl, walls, precision)


def cpu_baseline(case, st, steps: int, threads: int, single_phase: bool = False):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_backend import OracleSimulation, SPOracleSimulation, set_threads
    used = set_threads(threads)
    sim = (SPOracleSimulation if single_phase else OracleSimulation)(case, state=st)
    sim.run(1, 2)   # first touch
    t0 = time.perf_counter()
    sim.run(3, steps)
    dt = time.perf_counter() - t0
    sim.close()
    return case.n_fluid * steps / dt / 1e6, used


def traffic_from_profiles(name: str):
    p = os.path.join(REPO, "profiles", f"traffic_{name}.json")
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def timed_run(sim, args, nwhich, develop=0):
    """`develop` untimed steps, warm-up, timed steps, then a second pass of the same length with HIP-event kernel
    timing."""
    import ctypes as C
    ntime = 1
    if develop:
        sim.run(ntime, develop)
        ntime += develop
    sim.run(ntime, args.warmup)
    ntime += args.warmup
    sim.synchronize()
    t1 = time.perf_counter()
    sim.run(ntime, args.steps)
    sim.synchronize()
    dt = time.perf_counter() - t1
    ntime += args.steps
    sim._call("timing_enable", 1)
    sim.run(ntime, args.steps)
    means = []
    for which in nwhich:
        m, cnt = C.c_double(), C.c_int64()
        sim._call("timing", which, C.byref(m), C.byref(cnt))
        means.append(m.value)
    sim._call("timing_enable", 0)
    return dt, means, ntime + args.steps


OMP_BINDINGS = {"spread": {"OMP_PLACES": "cores", "OMP_PROC_BIND": "spread"}, "unbound": {}}


def host_description() -> dict:
    import platform
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "machine": platform.machine()}
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in txt.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[key.strip()] = val.strip()
    except OSError:
        pass
    return info


def cpu_child(args) -> None:
    """The CPU leg's measurement, in a process of its own so that its OpenMP runtime starts with the binding
    given in its environment (the parent's runtime may already be up)."""
    single = args.workload == "singlephase"
    case = workload_sp(args.size, "fp64") if single else workload(args.size, "fp64")
    threads = int(os.environ["OMP_NUM_THREADS"])
    v, used = cpu_baseline(case, case.initial_state(), args.cpu_steps, threads, single)
    print(json.dumps({"mlups": v, "threads": used}))


def cpu_leg(args, single_phase=False):
    """The oracle on the host cores: every core this process may use (OMP_NUM_THREADS when the environment
    sets it, as the GPU machines do for their CPU share), once with OMP_PLACES=cores OMP_PROC_BIND=spread and
    once without a binding; the faster is reported, both are in the sample text."""
    if args.no_cpu:
        return None
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    runs = {}
    for name, env_add in OMP_BINDINGS.items():
        env = {k: v for k, v in os.environ.items() if k not in ("OMP_PLACES", "OMP_PROC_BIND")}
        env.update(env_add, OMP_NUM_THREADS=str(threads))
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--workload", args.workload, "--size",
               str(args.size), "--cpu-steps", str(args.cpu_steps)]
        r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CPU leg ({name}) failed with status {r.returncode}")
        runs[name] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"[bench] cpu leg {name}: {runs[name]['mlups']:.2f} MLUPS on {runs[name]['threads']} threads",
              file=sys.stderr, flush=True)
    best = max(runs, key=lambda k: runs[k]["mlups"])
    host = host_description()
    return {"value": round(runs[best]["mlups"], 3), "unit": "MLUPS", "cores": runs[best]["threads"], "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {args.size}^3 workload (FP64, from its initial state) after "
                      f"2 untimed steps, C++/OpenMP oracle on {runs[best]['threads']} threads; "
                      + ", ".join(f"{k} {runs[k]['mlups']:.2f} MLUPS" for k in runs)
                      + f" ({OMP_BINDINGS['spread']['OMP_PLACES']}/{OMP_BINDINGS['spread']['OMP_PROC_BIND']} vs "
                      "no binding); host " + json.dumps(host, sort_keys=True),
            "binding": best, "host": host}


def roofline(b_alg, n_fluid, kernel_ms, traffic, kernel, **extra):
    achieved = b_alg * n_fluid / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    r = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic, "kernel": kernel,
         "kernel_ms": round(kernel_ms, 4)}
    r.update(extra)
    return r


def run_twophase(args) -> dict:
    from mflbm import Simulation
    t0 = time.time()
    case = workload(args.size, args.precision)
    st = case.initial_state()
    print(f"[bench] set-up {time.time() - t0:.1f}s, fluid nodes {case.n_fluid}", file=sys.stderr, flush=True)
    sim = Simulation(case, state=st)
    dt, means, ntime = timed_run(sim, args, (0, 1, 2), develop=args.develop)
    mon = sim.monitor()
    developed = None
    if args.developed_at > ntime:
        # the same run continued to a developed interface, then the same number of steps timed again
        sim.run(ntime, args.developed_at - ntime)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.run(args.developed_at, args.steps)
        sim.synchronize()
        dtd = time.perf_counter() - t1
        md = sim.monitor()
        developed = {"timed_from_step": args.developed_at, "value": round(case.n_fluid * args.steps / dtd / 1e6, 3),
                     "ms_per_step": round(dtd / args.steps * 1e3, 4), "saturation": md.saturation}
    sim.close()
    mlups = case.n_fluid * args.steps / dt / 1e6
    fp32 = args.precision != "fp64"
    return {
        "metric": METRIC if not fp32 else f"MLUPS (two-phase D3Q19 colour model, {args.precision} storage) at 256^3 porous",
        "value": round(mlups, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if fp32 else "f64", "data": "synthetic",
        "config": {"workload": f"two-phase drainage, {args.size}^3 seeded sphere pack porosity 0.22, 8-layer "
                               "buffers, velocity inlet 1e-4 + convective outlet, 150 deg through the invading "
                               "fluid (theta 30), "
                               + ({"fp32": "FP32 population storage", "mixed": "FP32 storage, FP64 collision"}
                                  .get(args.precision, "FP64")) + "; the reference's "
                               "colour model (two D3Q19 population sets: the reference has no D3Q7 lattice)",
                   "fluid_nodes": case.n_fluid, "lattice": [args.size] * 3,
                   "timed_from_step": args.develop + args.warmup + 1},
        "roofline": roofline(B_ALG32 if fp32 else B_ALG, case.n_fluid, means[0],
                             traffic_from_profiles(f"{'fp32_' if fp32 else ''}sphere_pack_{args.size}"),
                             "k_collide (even/odd mean)", gradient_chain_ms=round(means[1], 4),
                             step_ms_events=round(means[2], 4)),
        "developed_interface": developed,
        "cpu_baseline": cpu_leg(args),
        "check": {"saturation": mon.saturation, "umax": mon.umax, "failed": mon.failed},
    }


def run_singlephase(args) -> dict:
    from mflbm.singlephase import SPSimulation
    t0 = time.time()
    case = workload_sp(args.size, args.precision)
    st = case.initial_state()
    print(f"[bench] set-up {time.time() - t0:.1f}s, fluid nodes {case.n_fluid}", file=sys.stderr, flush=True)
    sim = SPSimulation(case, state=st)
    dt, means, _ = timed_run(sim, args, (0, 2))
    mon = sim.monitor()
    sim.close()
    mlups = case.n_fluid * args.steps / dt / 1e6
    fp32 = args.precision != "fp64"
    tag = {"fp64": "FP64", "fp32": "FP32 deviation storage", "mixed": "FP32 deviation storage, FP64 collision"}[args.precision]
    return {
        "metric": f"MLUPS (single-phase D3Q19 MRT, {tag}) at 256^3 porous",
        "value": round(mlups, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if fp32 else "f64", "data": "synthetic",
        "config": {"workload": f"single-phase permeability run, {args.size}^3 seeded sphere pack porosity 0.22, "
                               f"z-periodic, body force 1e-6, nu 0.1, {tag}",
                   "fluid_nodes": case.n_fluid, "lattice": [args.size] * 3},
        "roofline": roofline(B_ALG_SP32 if fp32 else B_ALG_SP, case.n_fluid, means[0],
                             traffic_from_profiles(f"sp{'32' if fp32 else ''}_sphere_pack_{args.size}"),
                             "k_sp_collide (even/odd mean)", step_ms_events=round(means[1], 4)),
        "cpu_baseline": cpu_leg(args, single_phase=True),
        "check": {"flowrate": mon.flowrate, "umax": mon.umax, "failed": mon.failed},
    }


def run_batch(args) -> dict:
    """96 REV cubes (sides from {48, 64, 96, 128}, sphere packs of porosity 0.15-0.35, seed 2024), nu 0.1,
    body force 1e-6, FP32 deviation storage with autotuned tile rows and block size.  value: MLUPS of the
    batched lattice update over all samples (all running); time_to_converged: wall time of the steady run
    until every sample meets the flow-rate criterion 1e-7 (or args.batch_max_steps): FP64 from the reference's
    start (the plain path), FP32 storage from the same start, and FP32 storage from the collision's rest state
    (rest_state = 1), with the FP32 permeability errors against FP64 over the samples converged in both."""
    import numpy as np
    from mflbm import batch
    t0 = time.time()
    samples = batch.rev_samples(args.batch_samples, seed=2024)
    print(f"[bench] {len(samples)} samples generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    p32 = args.precision if args.precision != "fp64" else "fp32"
    tune = batch.autotune(samples, p32)
    print(f"[bench] autotune: tile_rows {tune['tile_rows']} block {tune['block']}", file=sys.stderr, flush=True)
    kw = dict(la_nu=0.1, force_z=1e-6, ntime_monitor=500, tol=1e-7, tile_rows=tune["tile_rows"], block=tune["block"])
    b = batch.REVBatch(samples, precision=p32, **kw)
    b.run(1, args.warmup)
    b.synchronize()
    t1 = time.perf_counter()
    b.run(1 + args.warmup, args.steps)
    b.synchronize()
    dt = time.perf_counter() - t1
    nodes, pop_bytes = b.fluid_nodes, b.pop_bytes
    b.close()
    mlups = nodes * args.steps / dt / 1e6
    conv = {}
    for name, prec, rest in (("fp64", "fp64", 0), (p32, p32, 0), (p32 + "_rest_state", p32, 1)):
        bb = batch.REVBatch(samples, precision=prec, rest_state=rest, **kw)
        r = bb.run_steady(args.batch_max_steps, log=lambda m: print(f"[bench] {name} {m}", file=sys.stderr, flush=True))
        bb.close()
        conv[name] = r
        print(f"[bench] {name}: steady in {r.seconds:.2f}s, {int((r.conv_step > 0).sum())}/{len(samples)} converged, "
              f"max step {int(r.conv_step.max())}", file=sys.stderr, flush=True)
    errs = {}
    for name in (p32, p32 + "_rest_state"):
        both = (conv[name].conv_step > 0) & (conv["fp64"].conv_step > 0)
        rel = np.abs(conv[name].permeability / conv["fp64"].permeability - 1.0)[both]
        errs[name] = {"samples": int(both.sum()), "max": float(rel.max()) if both.any() else None,
                      "median": float(np.median(rel)) if both.any() else None}
    ms = dt / args.steps * 1e3
    return {
        "metric": f"MLUPS (batched single-phase REVs, D3Q19 MRT, {p32} FP32 deviation storage); time-to-converged permeability",
        "value": round(mlups, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(samples)} REV cubes, sides from {{48,64,96,128}}, sphere packs porosity "
                               "0.15-0.35 (seed 2024), z-periodic, body force 1e-6, nu 0.1, FP32 deviation storage, "
                               "one lattice-update launch per step for all samples",
                   "fluid_nodes": nodes, "samples": len(samples), "autotune": {k: tune[k] for k in ("tile_rows", "block")}},
        "roofline": roofline(B_ALG_SP32, nodes, ms, None, "k_spb_collide + periodic exchange (wall time per step)"),
        "time_to_converged": {
            p: {"seconds": round(conv[p].seconds, 3), "converged": int((conv[p].conv_step > 0).sum()),
                "max_step": int(conv[p].conv_step.max())} for p in conv},
        "permeability_error_vs_fp64": errs,
        "autotune_timings": tune["timings"],
        "cpu_baseline": None,
    }


def run_fp32_report(args) -> dict:
    """FP32 storage against FP64 on the 256^3 drainage (saturation, capillary number and the largest velocity at
    the report steps) and on a static drop (spurious-current magnitude: the largest velocity of a drop at rest,
    64^3, radius 16, 90 degrees, after args.drop_steps steps)."""
    from mflbm import Case, Simulation
    report = [int(x) for x in args.report_steps.split(",")]
    out = {"drainage": {}, "drop": {}, "ms_per_step": {}}
    precs = ("fp64", "fp32", "mixed")
    for prec in precs:
        case = workload(args.size, prec)
        sim = Simulation(case, device_init=True)
        t, rows = 0, []
        t0 = time.perf_counter()
        for n in report:
            sim.run(t + 1, n - t)
            t = n
            m = sim.monitor()
            rows.append({"step": n, "saturation": m.saturation, "ca": m.ca, "umax": m.umax})
            print(f"[bench] {prec} step {n}: sat {m.saturation:.8f} ca {m.ca:.6e} umax {m.umax:.6e}",
                  file=sys.stderr, flush=True)
        out["ms_per_step"][prec] = (time.perf_counter() - t0) / report[-1] * 1e3
        out["drainage"][prec] = rows
        sim.close()
        ctl = {"lattice_dimensions": (64, 64, 64), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
               "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 4, 4), "fluid1_viscosity": 1.0 / 6.0,
               "fluid2_viscosity": 1.0 / 6.0, "surface_tension": 0.01, "theta": 90.0, "body_force_0": 0.0,
               "initial_fluid_distribution_option": 5, "initial_interface_position": 16.0}
        import numpy as np
        drop = Case(ctl, np.zeros((64, 64, 64), np.int8), precision=prec)
        sim = Simulation(drop, device_init=True)
        sim.run(1, args.drop_steps)
        m = sim.monitor()
        out["drop"][prec] = {"steps": args.drop_steps, "umax": m.umax, "saturation": m.saturation}
        sim.close()
    err = {}
    for prec in precs[1:]:
        err[prec] = []
        for a, b in zip(out["drainage"][prec], out["drainage"]["fp64"]):
            err[prec].append({"step": a["step"], "saturation_abs": abs(a["saturation"] - b["saturation"]),
                              "ca_rel": abs(a["ca"] / b["ca"] - 1.0) if b["ca"] else None,
                              "umax_rel": abs(a["umax"] / b["umax"] - 1.0) if b["umax"] else None})
    d64 = out["drop"]["fp64"]["umax"]
    spur = {"fp64": d64}
    for prec in precs[1:]:
        d = out["drop"][prec]["umax"]
        spur[prec] = {"umax": d, "rel": abs(d / d64 - 1.0) if d64 else None}
    return {"metric": "FP32 storage vs FP64: two-phase 256^3 drainage and static drop", "unit": "report",
            "value": None, "n_gpus": 1, "data": "synthetic", "vs_fp64": err, "spurious_current": spur,
            "detail": out}


def run_warmstart(args) -> dict:
    """Learned warm start: steady fields of args.ws_samples seeded 64^3 packs from the batched solver, the
    encoder-decoder trained on them for args.ws_epochs epochs with the project's kernels, then unseen packs of
    sides args.ws_eval run to the flow-rate criterion 1e-7 three ways: the plain path (FP64, the reference's
    initial state), from rest at the collision's equilibrium and from the network's fields (both with FP32
    storage and FP32 collision).  ``value`` is the warm-started wall time over the plain path's."""
    import numpy as np
    from mflbm import batch, warmstart
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    data = warmstart.make_dataset(args.ws_samples, side=64, seed=11, log=log)
    log(f"dataset: {args.ws_samples} fields in {data.seconds:.1f}s, steps median {int(np.median(data.conv_steps))}, "
        f"target rms (u, v, w, p) {np.round(data.target_rms, 4).tolist()}")
    net = warmstart.WarmStartNet(seed=0)
    t1 = time.perf_counter()
    hist = warmstart.train(net, data, epochs=args.ws_epochs, batch_size=4, lr=1e-3, log=log)
    import torch
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    iters = args.ws_epochs * (args.ws_samples // 4)
    results = []
    for k, side in enumerate(int(v) for v in args.ws_eval.split(",")):
        s = batch.rev_samples(1, seed=90001 + k, sides=(side,), porosity=(0.25, 0.25))[0]
        r = warmstart.compare_runs(net, s)
        row = {"side": side, "predict_s": round(r.seconds_predict, 3), "k_rel_err_vs_plain": r.k_rel_err,
               "time_ratio_vs_plain": round(r.time_ratio, 4)}
        for name in ("plain", "cold", "warm"):
            rec = getattr(r, name)
            row[name] = {"steps": rec.steps, "seconds": round(rec.seconds, 3), "permeability": rec.permeability,
                         "status": rec.status}
        log(f"{side}^3: steps plain {r.plain.steps} cold {r.cold.steps} warm {r.warm.steps}; time plain "
            f"{r.plain.seconds:.2f}s cold {r.cold.seconds:.2f}s warm {r.warm.seconds:.2f}s (ratio {r.time_ratio:.3f}, "
            f"probe + prediction {r.seconds_predict:.2f}s); k rel err {r.k_rel_err:.2e}")
        results.append(row)
    return {"metric": "time-to-converged permeability: warm-started FP32-storage run over the plain FP64 run",
            "unit": "ratio", "value": results[0]["time_ratio_vs_plain"] if results else None,
            "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
            "config": {"workload": f"{args.ws_samples} training fields (64^3 sphere packs, porosity 0.15-0.35, "
                                   f"batched solver, criterion 1e-6), {args.ws_epochs} epochs, BF16 encoder-decoder "
                                   "(4 levels, 32 base channels); unseen packs (porosity 0.25) to criterion 1e-7: "
                                   "plain = FP64 from the reference's initial state, cold = FP32 storage + FP32 "
                                   "collision from rest at the collision's equilibrium, warm = the same for 200 "
                                   "steps, then restarted from the network's (u, v, w, rho)"},
            "training": {"data_seconds": round(data.seconds, 1), "train_seconds": round(t_train, 1),
                         "target_rms": [round(float(v), 5) for v in data.target_rms],
                         "iterations": iters, "loss_first_epoch": hist[0] if hist else None,
                         "loss_last_epoch": hist[-1] if hist else None},
            "results": results, "total_seconds": round(time.perf_counter() - t0, 1)}


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix-core peak


def run_nn(args) -> dict:
    """Matrix-core throughput of the warm-start network's convolution kernels: the forward convolution
    (mflbm_nn_conv3d_fwd, implicit GEMM on v_mfma_f32_16x16x32_bf16) and the fixed-order weight gradient
    (mflbm_nn_conv3d_dw_ordered) on the layer shapes of the network's levels at the training shape (4 fields of
    64^3: level l is 64 / 2^l per side with 32 * 2^l channels).  2 * positions * 27 * cin * cout flops per call,
    timed with HIP events on the stream the kernels run on; ``value`` is the forward kernel's TFLOP/s on the
    first level."""
    import ctypes as C
    import numpy as np
    import torch
    from mflbm import native
    lib = native.load()
    P = lambda t: C.c_void_p(t.data_ptr())
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    rows = []
    for lvl, (side, ch) in enumerate(((64, 32), (32, 64), (16, 128), (8, 256))):
        n, cin, cout = 4, ch, ch
        shape = (n, side, side, side)
        M = n * side ** 3
        kpad = (27 * cin + 31) // 32 * 32
        x = (torch.randn(M * cin, device="cuda") * 0.5).to(torch.bfloat16)
        g = (torch.randn(M * cout, device="cuda") * 0.5).to(torch.bfloat16)
        w = torch.randn(cout * 27 * cin, device="cuda") * 0.05
        wf = torch.zeros(cout * kpad, dtype=torch.bfloat16, device="cuda")
        wb = torch.zeros(cin * ((27 * cout + 31) // 32 * 32), dtype=torch.bfloat16, device="cuda")
        bias = torch.zeros(cout, device="cuda")
        y = torch.empty(M * cout, dtype=torch.bfloat16, device="cuda")
        dw = torch.zeros(cout * 27 * cin, device="cuda")
        db = torch.zeros(cout, device="cuda")
        need = int(lib.mflbm_nn_conv3d_dw_work(*shape, cin, cout))
        work = torch.empty(need, device="cuda")
        assert lib.mflbm_nn_weights_to_bf16(P(w), P(wf), P(wb), cin, cout, sp) == 0
        flops = 2.0 * M * 27 * cin * cout
        out = {"level": lvl, "positions": M, "cin": cin, "cout": cout, "gflop_per_call": round(flops / 1e9, 2)}
        for name, call in (("fwd", lambda: lib.mflbm_nn_conv3d_fwd(P(x), P(wf), P(bias), P(y), *shape, cin, cout, 1, sp)),
                           ("dw", lambda: lib.mflbm_nn_conv3d_dw_ordered(P(x), P(g), P(dw), P(db), P(work), need, *shape,
                                                                         cin, cout, sp))):
            for _ in range(3):
                assert call() == 0
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 10
            e0.record(st)
            for _ in range(reps):
                call()
            e1.record(st)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            tf = flops / (ms * 1e-3) / 1e12
            out[name] = {"ms": round(ms, 4), "tflops": round(tf, 2), "frac_of_bf16_peak": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}
        rows.append(out)
        print(f"[bench] nn level {lvl}: {out}", file=sys.stderr, flush=True)
    return {"metric": "warm-start convolution throughput (BF16 matrix cores)", "value": rows[0]["fwd"]["tflops"],
            "unit": "TFLOP/s", "n_gpus": 1, "dtype": "bf16", "data": "synthetic",
            "roofline": {"bound": "mfma", "achieved": rows[0]["fwd"]["tflops"], "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": rows[0]["fwd"]["frac_of_bf16_peak"], "traffic": None,
                         "kernel": "k_conv_fwd (level 0: 4 x 64^3, 32 -> 32 channels)"},
            "layers": rows}


def run_copy(args) -> dict:
    """Calibration workload for the hardware byte counters: a plain device-to-device copy of args.copy_mib MiB.
    Under ``rocprofv3 --pmc FETCH_SIZE`` (or ``WRITE_SIZE``) the copy kernel reads (writes) exactly that many
    bytes per dispatch; the ratio of counter to size is the factor that turns the solver kernels' counters into
    HBM bytes.  ``value`` is the copy's read + write bandwidth, the practical ceiling of a streaming kernel."""
    import torch
    n = args.copy_mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    reps = 5
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    gbs = 2.0 * args.copy_mib * (1 << 20) / (ms * 1e-3) / 1e9
    return {"metric": "device copy bandwidth (read + write)", "value": round(gbs, 1), "unit": "GB/s",
            "n_gpus": 1, "ms_per_copy": round(ms, 4), "bytes_each_way": args.copy_mib * (1 << 20),
            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "data": "synthetic"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("twophase", "singlephase", "batch", "fp32report", "warmstart", "copy", "nn"),
                    default="twophase")
    ap.add_argument("--copy-mib", type=int, default=1024, help="copy workload: MiB copied per dispatch")
    ap.add_argument("--batch-samples", type=int, default=96)
    ap.add_argument("--report-steps", default="5000,10000,20000")
    ap.add_argument("--ws-samples", type=int, default=256, help="warm start: training fields (the full set is 2000)")
    ap.add_argument("--ws-epochs", type=int, default=40)
    ap.add_argument("--ws-eval", default="128", help="warm start: sides of the unseen evaluation packs")
    ap.add_argument("--drop-steps", type=int, default=4000)
    ap.add_argument("--batch-max-steps", type=int, default=100000)
    ap.add_argument("--precision", choices=("fp64", "fp32", "mixed"), default="fp64",
                    help="population storage: fp64 (reference arithmetic) or fp32")
    ap.add_argument("--develop", type=int, default=DEVELOP_STEPS,
                    help="two-phase: untimed steps before the warm-up, so that the timed steps see a developed "
                         "interface")
    ap.add_argument("--developed-at", type=int, default=3000,
                    help="two-phase: after the measurement, continue the run to this step and time the same "
                         "number of steps again (0: off)")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs a single domain on one MI355X (multi-GPU z-slab "
              "decomposition is a design, not part of the build); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    run = {"singlephase": run_singlephase, "batch": run_batch, "fp32report": run_fp32_report,
           "warmstart": run_warmstart, "copy": run_copy, "nn": run_nn}.get(args.workload, run_twophase)
    out = run(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
