#!/usr/bin/env python3
"""Headline benchmark: object-steps/s of the propagate + conjunction hot path.

Workload (one GPU): 1,048,576 objects of the debris-tail population, full force
model (Kepler, J2, C22, S22, Sun, Moon, SRP, drag), dt = 10 s, linked-cell
conjunction scan every step with a 10 km cutoff and a 1 km absolute threshold.
A step is one iteration of the simulation loop: Yoshida step of every object,
burn-up below 150 km (Simulation::deleteBurnUps), cell binning, pair scan with
the exact FP64 closest-approach test.

Both precision forms run the same steps from the same initial state:
  value          object-steps/s of the split form (FP32 perturbations and prefilter)
  fp64_value     the plain FP64 form
  pos_err_cm_*   split vs FP64 position difference after the timed steps
  pairs_equal    whether both forms report the same (iteration, id pair) set

Inputs are resident in HBM before the timed region; the timed region is
bracketed by stream synchronisations.  The roofline entry covers the kernel
with the largest device time in a profiled pass (HIP events around each launch
on the engine's stream).  The CPU baseline times the C oracle (restatement of
the reference path, OpenMP) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "object-steps/s @1M objects: split-prec vs FP64 GPU at ≤1 cm/day pos error"
MIN_ALT = 150.0  # sim/minAltitude default (cfg/default_cfg.yaml)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic HBM bytes per object and launch (per object-step for propagate),
# counted from the kernels' own loads and stores.
BYTES = {
    # r, v read + write (96) + srp + drag coefficients (16 FP64 | 4 FP32 + 8 FP64) + activity (4) + cell key (4)
    "propagate": {"fp64": 96 + 16 + 4 + 4, "split": 96 + 12 + 4 + 4},
    # keys read (4) + r read (24) + record write (32 | 16)
    "tile_partition": {"fp64": 4 + 24 + 32, "split": 4 + 24 + 16},
    # every record read once for its own plane and once as the next plane of the plane below
    "pair_tiles": {"fp64": 2 * 32, "split": 2 * 16},
    # r read (24) + slot read (4) + bucket start (4) + record write (32 | 16) + index write (4)
    "bin_scatter": {"fp64": 24 + 4 + 4 + 32 + 4, "split": 24 + 4 + 4 + 16 + 4},
    # own record (32 | 16) + own index (4) + 4 run bounds (32); neighbour records are hits in L2
    "pair_scan": {"fp64": 32 + 4 + 32, "split": 16 + 4 + 32},
}


# HBM bytes per launch at 1,048,576 objects (read + write), measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by a 256 MiB copy kernel, on the plane-tile shape the
# autotuner picks for this workload (FP64: 512-thread pair workgroups, 4,096-object tiles;
# split: 256 threads, 8,192-object tiles; both: buckets of 2 x 2 cells); the runs and the
# commands are in
# profiles/hbm_traffic_1M_full_model.md.
# FP64 and FP32 VALU flops per object-step of the propagation kernel (full force model),
# from rocprofv3 SQ_INSTS_VALU_FLOPS_{FP64,FP32} (+ the _TRANS counters), per wavefront
# instruction and lane, on the 1M-object workload.
MEASURED_FLOPS_PER_OBJECT = {
    ("propagate", "fp64"): (794.0 + 15.0, 0.0),
    ("propagate", "split"): (551.0 + 6.0, 249.0 + 12.0),
}
FP64_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (non-packed issue is half of this)

MEASURED_TRAFFIC_1M = {
    ("propagate", "fp64"): (71.47 + 54.65) * 1e6,
    ("propagate", "split"): (67.27 + 54.65) * 1e6,
    ("tile_partition", "fp64"): (29.43 + 35.93) * 1e6,
    ("tile_partition", "split"): (29.45 + 17.83) * 1e6,
    ("pair_tiles", "fp64"): (56.35 + 0.06) * 1e6,
    ("pair_tiles", "split"): (32.00 + 0.06) * 1e6,
}


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=8640, help="timed steps (8640 = one day of 10 s steps)")
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--objects", type=int, default=1 << 20)
    p.add_argument("--profile-steps", type=int, default=500)
    p.add_argument("--cpu-steps", type=int, default=60, help="steps of the CPU baseline sample (~15 s on 16 threads)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-propagation leg")
    p.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    p.add_argument("--ensemble-steps", type=int, default=2000)
    p.add_argument("--no-clustered", action="store_true", help="skip the clustered 8M-object leg")
    p.add_argument("--clustered-steps", type=int, default=10,
                   help="timed steps of the clustered leg (each ~0.4 s: the workload has ~2.7e9 conjunctions "
                        "per step)")
    p.add_argument("--no-dropin", action="store_true", help="skip the drop-in adapter leg")
    p.add_argument("--dropin-steps", type=int, default=50)
    p.add_argument("--forms", default="fp64,split")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default plane-tile shape (profiler runs: the tuner's timings change under "
                        "counter collection)")
    p.add_argument("--tuning", metavar="JSON",
                   help="apply these plane-tile shapes instead of autotuning: a JSON object per form, as the "
                        "'tuning' field of an earlier run's line (profiler runs then measure the tuned shape)")
    p.add_argument("--ensemble-tuning", metavar="JSON",
                   help="apply this plane-tile shape (one JSON object) in the ensemble leg instead of autotuning it")
    p.add_argument("--clustered-tuning", metavar="JSON",
                   help="apply this plane-tile shape (one JSON object, both forms) in the clustered leg instead "
                        "of autotuning it")
    p.add_argument("--profile-mode", choices=("overlapped", "serial"), default="overlapped",
                   help="profiled pass: the overlapped loop (default), or one kernel at a time (propagate one "
                        "step, then scan its pairs)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="first copy MIB MiB device-to-device 3 times (a plain copy kernel of known size, "
                        "for scaling hardware byte counters), then run the bench")
    return p.parse_args()


def calibrate_copy(mib):
    import torch

    x = torch.ones(mib << 20, dtype=torch.uint8, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"# copied {mib} MiB x 3 (device to device)", file=sys.stderr, flush=True)


def make_engine(la, precision):
    # the tuner times each shape over simulate() steps (the loop this bench times)
    cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 24,
                          tune_simulate=1)
    return la.Engine(cfg)


def run_form(la, pop, precision, steps, warmup, profile_steps, autotune=True, pinned=None, serial=False):
    eng = make_engine(la, precision)
    eng.set_objects(pop)
    # plane-tile shape picked by timing the scan on the initial state (outside the timed region),
    # or a shape an earlier run picked
    if pinned:
        eng.set_tuning(pinned)
        tuning = dict(pinned)
    else:
        # best of 5 bursts per shape: shapes within ~2% of each other in step time are ranked
        # on less noise (tuning runs before the timed region)
        tuning = eng.autotune(repetitions=5) if autotune else None
    if warmup:
        eng.simulate(warmup, first_iteration=0, tpcd=1, min_altitude_km=MIN_ALT)
    eng.clear_log()
    eng.set_objects(pop)
    eng.set_time(0.0)
    eng.synchronize()
    t0 = time.perf_counter()
    stats = eng.simulate(steps, first_iteration=0, tpcd=1, min_altitude_km=MIN_ALT)
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    r, _ = eng.state()
    ids = eng.ids()
    log = eng.fetch_log()
    # profiled pass: per-kernel device time from HIP events on the engine's stream
    eng.clear_log()
    eng.set_objects(pop)
    eng.set_time(0.0)
    eng.set_profiling(True)
    if serial:
        # one kernel at a time: propagation step, then the scan of its state (same kernels as
        # the loop's, without the other stream's work beside them)
        for _ in range(profile_steps):
            eng.propagate(1)
            eng.conjunctions()
    else:
        eng.simulate(profile_steps, first_iteration=0, tpcd=1, min_altitude_km=MIN_ALT)
    kernels = {}
    for name in ("propagate", "bin_keys", "tile_partition", "pair_tiles", "bin_reduce", "bin_scan", "bin_scatter",
                 "pair_scan"):
        launches, ms = eng.kernel_stats(name)
        if launches:
            kernels[name] = {"launches": launches, "avg_us": 1000.0 * ms / launches}
    eng.set_profiling(False)
    eng.close()
    return {"elapsed": elapsed, "stats": stats, "r": r, "ids": ids, "log": log, "kernels": kernels, "tuning": tuning}


def cpu_baseline(pop, steps):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol

    threads = min(16, len(os.sched_getaffinity(0)))
    cfg = ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0)

    def run(n, nsteps, nthreads):
        sel = slice(0, n)
        r, v = pop.r()[sel], pop.v()[sel]
        t0 = time.perf_counter()
        for s in range(nsteps):
            r, v = ol.propagate(r, v, s * 10.0, 10.0, 0xFF, 1, pop.aom[sel], pop.bc_inv[sel], pop.cr[sel],
                                pop.activity[sel], nthreads)
            ol.pairs(ol.ObjectSet(r, v, pop.radius_m[sel], pop.activity[sel], pop.id[sel], pop.parent[sel]), cfg,
                     "grid", nthreads)
        return time.perf_counter() - t0

    n = pop.n
    el = run(n, steps, threads)
    n1, steps1 = min(n, 1 << 17), 5
    el1 = run(n1, steps1, 1)
    model = "unknown"
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": n * steps / el, "unit": "object-steps/s", "cores": threads, "kind": "port",
            "sample": f"C oracle (OpenMP, -O3 -march=x86-64-v3, no FMA contraction): {n} objects x {steps} steps of the same "
                      f"workload (propagate + linked-cell scan), {el:.1f} s on {threads} host threads",
            "cpu_model": model, "nproc": os.cpu_count(), "threads_available": len(os.sched_getaffinity(0)),
            "single_thread": {"value": n1 * steps1 / el1, "sample": f"{n1} objects x {steps1} steps, {el1:.1f} s"}}


def adjoint_leg(la, reps=3):
    """Differentiable propagation: 4,096 objects of the 16,384-object workload's distributions,
    full force model in FP64, dt = 60 s, 1,440 steps (one day), checkpoints every 64 steps;
    loss = sum |r_T - target|^2 with targets = nominal final state + N(0, 100 m).  Reports
    gradient-steps/s (objects x steps per forward + backward pass, device time from events on
    torch's stream)."""
    import torch
    from ladds_amd import adjoint

    n, steps, dt = 4096, 1440, 60.0
    pop = la.generate_population(0, 0x1ADD5 + 4, n)
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    plan = adjoint.AdjointPlan(n, steps, dt, components=la.ALL_COMPONENTS, checkpoint_every=64)
    aom_t, cr_t = t(pop.aom), t(pop.cr)
    with torch.no_grad():
        rT0, _ = adjoint.propagate(t(pop.r()), t(pop.v()), t(pop.bc_inv), plan, aom=aom_t, cr=cr_t)
    target = rT0 + 0.1 * torch.randn(rT0.shape, dtype=torch.float64, device=dev,
                                     generator=torch.Generator(device=dev).manual_seed(4))
    times = []
    for rep in range(reps + 1):
        r0, v0, bc = t(pop.r()).requires_grad_(), t(pop.v()).requires_grad_(), t(pop.bc_inv).requires_grad_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rT, _ = adjoint.propagate(r0, v0, bc, plan, aom=aom_t, cr=cr_t)
        loss = ((rT - target) ** 2).sum()
        loss.backward()
        e1.record()
        torch.cuda.synchronize()
        if rep:
            times.append(e0.elapsed_time(e1) * 1e-3)
    sec = float(np.median(times))
    plan.close()
    out = {"workload": f"{n} objects, full force model FP64, dt {dt:g} s, {steps} steps, checkpoints every 64",
           "gradient_steps_per_s": n * steps / sec, "ms_per_gradient": sec * 1e3}
    grads = {"pop": pop, "g_r": 2.0 * (rT.detach().cpu().numpy() - target.cpu().numpy()),
             "got": np.concatenate([r0.grad.cpu().numpy(), v0.grad.cpu().numpy()], 1),
             "got_bc": bc.grad.cpu().numpy(), "steps": steps, "dt": dt}
    return out, grads


def cpu_baseline_adjoint(la, grads):
    """CPU leg of the differentiable propagation: the oracle's FP64 complex-step gradient of
    the same loss, timed, and the GPU gradient's relative error against it."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol

    pop, g_r = grads["pop"], grads["g_r"]
    threads = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    cr_, cv_, cbc = ol.vjp(pop.r(), pop.v(), 0.0, grads["dt"], la.ALL_COMPONENTS, grads["steps"], g_r,
                           np.zeros_like(g_r), aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr, nthreads=threads)
    cpu_sec = time.perf_counter() - t0
    ref = np.concatenate([cr_, cv_], 1)
    rel = np.max(np.abs(grads["got"] - ref), 1) / np.max(np.abs(ref), 1)
    bc_rel = np.abs(grads["got_bc"] - cbc) / np.max(np.abs(cbc))
    return {"grad_rel_err_max": float(rel.max()), "grad_rel_err_median": float(np.median(rel)),
            "bc_grad_rel_err_max": float(bc_rel.max()),
            "cpu_reference": {"kind": "port", "method": "oracle complex-step FP64 gradient (7 complex propagations "
                                                        "per object)",
                              "gradient_steps_per_s": pop.n * grads["steps"] / cpu_sec, "cores": threads}}


def _pair_hash(a, b):
    """Order-independent per-pair hash (splitmix64 finaliser of the id pair), for set digests."""
    def mix(x):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))
    with np.errstate(over="ignore"):
        return mix(a ^ mix(b))


def ensemble_leg(la, steps, warmup=10, pinned=None):
    """Batched ensemble: 256 scenarios of log-uniform 1,024-65,536 objects (~4M in total), each
    a seeded population of the 16,384-object distributions; full force model, split
    precision, dt = 10 s, detection every step (10 km cutoff, 1 km threshold), burn-up below
    150 km.  Each step is one replay of a captured HIP graph with per-scenario cell grids."""
    rng = np.random.default_rng(0x1ADD5 + 3)
    sizes = np.exp(rng.uniform(np.log(1024), np.log(65536), 256)).astype(np.int64)
    pop, off = la.generate_ensemble(sizes, (0x1ADD5 + 3) << 8, workload=0)
    out = {"workload": f"{len(sizes)} scenarios, {pop.n} objects (log-uniform 1024-65536 per scenario), full "
                       f"force model, split precision, dt 10 s, {steps} steps, detection every step",
           "objects": int(pop.n), "steps": steps}
    res = {}
    # tuned on simulate bursts, like the 1M legs: single detections cannot tell apart shapes
    # that differ by ~7% in the overlapped loop here (128 vs 512 buckets per plane)
    cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=la.SPLIT, cutoff_km=10.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 22,
                          tune_simulate=1)
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(off)
        if pinned:
            eng.set_tuning(pinned)
            tuning = dict(pinned)
        else:
            tuning = eng.autotune(repetitions=2)
        eng.simulate(warmup, tpcd=1, min_altitude_km=MIN_ALT)
        for graph in (1, 0):  # the same tuned context, steps replayed from a graph or launched eagerly
            eng.configure(step_graph=graph)
            eng.clear_log()
            eng.set_objects(pop)
            eng.set_scenarios(off)
            eng.set_time(0.0)
            eng.synchronize()
            t0 = time.perf_counter()
            st = eng.simulate(steps, tpcd=1, min_altitude_km=MIN_ALT)
            eng.synchronize()
            el = time.perf_counter() - t0
            res[graph] = (el, st, eng.fetch_log(), tuning)
    el, st, log, tuning = res[1]
    scen = log["scenario"].astype(np.int64)
    h = _pair_hash(log["pair"]["id_a"], log["pair"]["id_b"] ^ (log["iteration"] << np.uint64(40)))
    counts = np.bincount(scen, minlength=len(sizes))
    digests = np.zeros(len(sizes), np.uint64)
    with np.errstate(over="ignore"):
        np.add.at(digests, scen, h)
    el0, st0, log0, _ = res[0]
    same = (st["collisions"] == st0["collisions"] and np.array_equal(np.sort(h), np.sort(
        _pair_hash(log0["pair"]["id_a"], log0["pair"]["id_b"] ^ (log0["iteration"] << np.uint64(40))))))
    out.update({"object_steps_per_s": pop.n * steps / el, "ms_per_step": 1e3 * el / steps,
                "eager_object_steps_per_s": pop.n * steps / el0, "graph_equals_eager": bool(same),
                "collisions": int(st["collisions"]), "evasions": int(st["evasions"]),
                "remaining": int(st["remaining"]), "tuning": tuning,
                "pairs_per_scenario": counts.tolist(),
                "pairset_digest_per_scenario": [f"{int(d):016x}" for d in digests]})
    return out


def clustered_cpu_baseline(la, pop, ids, r, v, gpu_digest):
    """CPU leg of the clustered workload: the oracle's digest-only grid scan (OpenMP on the
    host threads) of the state the GPU reached, timed, and whether its count and digest
    equal the GPU's for that state (a full-size parity check of ~2.7e9 conjunctions)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol

    threads = min(16, len(os.sched_getaffinity(0)))
    k = ids.astype(np.int64)  # ids are the generated indices; burnt-up objects are gone
    objs = ol.ObjectSet(r, v, pop.radius_m[k], pop.activity[k], pop.id[k], pop.parent[k])
    t0 = time.perf_counter()
    want = ol.grid_digest(objs, ol.functor_cfg(5.0, 10.0, threshold_mode=1, threshold_km=0.5), threads)
    el = time.perf_counter() - t0
    return {"kind": "port", "method": "oracle digest-only linked-cell scan of the GPU's final state",
            "objects": int(pop.n), "seconds": el, "cores": threads,
            "conjunctions_per_s": want[0] / el, "conjunctions": int(want[0]),
            "digest_equals_gpu": bool(want == gpu_digest)}


def clustered_leg(la, steps, cpu_check=True, pinned=None):
    """8,388,608 objects, 60 % replicated constellation slots (+-50 m) and 40 % background;
    split precision, dt = 10 s, cutoff 5 km, 0.5 km threshold, pair output as count + digest;
    the autotuner picks cell size and tile shape (or `pinned` applies one).  The FP64 and
    split forms' digests of the initial state are compared.  Timed over `steps` steps of the
    1,000-step configuration."""
    n = 8388608
    pop = la.generate_population(2, 0x1ADD5 + 2, n)
    out = {"workload": f"{n} objects clustered (constellation shells), split precision, dt 10 s, cutoff 5 km, "
                       f"threshold 0.5 km, digest output; {steps} of 1000 steps timed", "objects": n, "steps": steps}
    digests = {}
    for prec, name in ((la.FP64, "fp64"), (la.SPLIT, "split")):
        cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=prec, cutoff_km=5.0,
                              threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=0.5, pair_output=la.PAIRS_DIGEST,
                              log_capacity=1 << 16)  # a step takes ~75 ms: tune on single detections
        with la.Engine(cfg) as eng:
            eng.set_objects(pop)
            if pinned:
                eng.set_tuning(pinned)
                tuning = dict(pinned)
            else:
                tuning = eng.autotune(repetitions=1)
            digests[name] = eng.conjunction_digest(10.0)
            if name == "split":
                eng.synchronize()
                t0 = time.perf_counter()
                st = eng.simulate(steps, tpcd=1, min_altitude_km=MIN_ALT)
                eng.synchronize()
                el = time.perf_counter() - t0
                out.update({"object_steps_per_s": n * steps / el, "ms_per_step": 1e3 * el / steps,
                            "conjunctions_per_step": st["collisions"] / max(steps, 1), "tuning": tuning})
                if cpu_check:
                    final = eng.conjunction_digest(10.0)
                    r, v = eng.state()
                    out["cpu_baseline"] = clustered_cpu_baseline(la, pop, eng.ids(), r, v, final)
    out["initial_conjunctions"] = int(digests["split"][0])
    out["split_digest_equals_fp64"] = bool(digests["split"] == digests["fp64"])
    return out


def dropin_leg(steps):
    """The drop-in path as LADDS drives it: the C++ adapter (include/ladds_amd.hpp) built like
    Simulation.cpp:122-148, then per iteration YoshidaIntegrator::integrate(false) and a fresh
    CollisionFunctor through Container::iteratePairwise (Simulation.cpp:297, :211-222), with the
    reference's radius-scaled threshold.  Timed by tests/cpp/adapter_check's `timeloop`
    (steady clock around the loop, after a warm-up loop), FP64 form, on the 16,384-object
    Kepler + J2 workload and the 1,048,576-object full-model workload."""
    import subprocess
    import tempfile

    exe = ROOT / "tests" / "cpp" / "adapter_check"
    if not exe.exists():
        return {"error": f"{exe} not built"}
    out = {}
    for name, n, comps, workload in (("16k_kepler_j2", 16384, "1 1 0 0 0 0 0 0", 0),
                                     ("1m_full_model", 1 << 20, "1 1 1 1 1 1 1 1", 1)):
        scen = [f"container 10.0 {0} 0", f"accumulator {comps} 0.0", "integrator 10.0",
                f"population {workload} {0x1ADD5 + workload} {n} 0",
                "timeloop 5 10.0 10.0 1.0 0.05 0.1", f"timeloop {steps} 10.0 10.0 1.0 0.05 0.1"]
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write("\n".join(scen) + "\n")
            path = f.name
        res = subprocess.run([str(exe), path], capture_output=True, text=True, timeout=600)
        os.unlink(path)
        if res.returncode != 0:
            out[name] = {"error": res.stderr[-500:]}
            continue
        t = [ln.split() for ln in res.stdout.splitlines() if ln.startswith("timing ")][-1]
        sec = float(t[2])
        out[name] = {"objects": n, "steps": steps, "object_steps_per_s": n * steps / sec,
                     "ms_per_step": 1e3 * sec / steps, "collisions": int(t[3]), "evasions": int(t[4])}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the engine has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    dropin = None
    if not a.no_dropin:
        # first: the adapter's loop (its own process) is host-latency bound, and run after the
        # other legs its 16k step measured 0.132 ms against 0.060 ms run first (an idle GPU
        # context in this process beside it leaves it at 0.060 ms)
        dropin = dropin_leg(a.dropin_steps)
        print(f"# dropin: {json.dumps(dropin)}", file=sys.stderr, flush=True)
    if not a.no_adjoint or a.calibrate_copy:
        # torch ships its own HIP runtime: importing it first binds the engine library to it
        print("# importing torch", file=sys.stderr, flush=True)
        import torch  # noqa: F401
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
    import ladds_amd as la

    if la.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    t_gen = time.perf_counter()
    pop = la.generate_population(1, 0x1ADD5 + 1, a.objects)
    print(f"# population {a.objects} generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    forms = {"fp64": la.FP64, "split": la.SPLIT}
    res = {}
    text = (a.tuning or "{}").strip()
    pinned = json.loads(text if text.startswith("{") else Path(text).read_text())
    for name in a.forms.split(","):
        res[name] = run_form(la, pop, forms[name], a.steps, a.warmup, a.profile_steps, not a.no_autotune,
                             pinned.get(name), a.profile_mode == "serial")
        r = res[name]
        print(f"# {name}: {a.steps} steps in {r['elapsed']:.3f} s, {a.objects * a.steps / r['elapsed']:.4g} "
              f"object-steps/s, collisions {r['stats']['collisions']}, candidates {r['stats']['candidates']}, "
              f"kernels {json.dumps(r['kernels'])}", file=sys.stderr, flush=True)

    main_form = "split" if "split" in res else next(iter(res))
    m = res[main_form]
    value = a.objects * a.steps / m["elapsed"]
    out = {
        "metric": METRIC, "value": value, "unit": "object-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1000.0 * m["elapsed"] / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32" if main_form == "split" else "f64",
        "data": "synthetic (seeded debris-tail population)",
        "config": {"workload": f"{a.objects} objects, full force model, dt 10 s, {a.steps} steps, cutoff 10 km, "
                               f"1 km absolute threshold, detection every step",
                   "objects": a.objects, "precision": main_form},
    }
    if "fp64" in res and "split" in res:
        f = res["fp64"]
        out["fp64_value"] = a.objects * a.steps / f["elapsed"]
        out["split_over_fp64"] = out["value"] / out["fp64_value"]
        sp = res["split"]
        common, i_s, i_f = np.intersect1d(sp["ids"], f["ids"], return_indices=True)
        err = np.linalg.norm(sp["r"][i_s] - f["r"][i_f], axis=1) * 1e5
        out["pos_err_cm_max"] = float(err.max())
        out["pos_err_cm_median"] = float(np.median(err))
        out["remaining"] = {"split": int(sp["ids"].shape[0]), "fp64": int(f["ids"].shape[0]),
                            "common": int(common.shape[0])}

        def pairkeys(log):
            p = log["pair"]
            keys = np.stack([log["iteration"], p["id_a"], p["id_b"], log["evaded"].astype(np.uint64)], 1)
            return keys[np.lexsort(keys.T[::-1])]

        ks, kf = pairkeys(res["split"]["log"]), pairkeys(f["log"])
        out["pairs_equal"] = bool(ks.shape == kf.shape and np.array_equal(ks, kf))
        out["pairs"] = {"split": int(ks.shape[0]), "fp64": int(kf.shape[0])}
    # roofline of the dominant kernel of the headline form (and of every kernel, per form)
    def roof(form, name, k):
        per_obj = BYTES.get(name, {}).get(form)
        if not per_obj:
            return None
        achieved = per_obj * a.objects / (k["avg_us"] * 1e-6) / 1e9
        traffic = MEASURED_TRAFFIC_1M.get((name, form)) if a.objects == 1 << 20 else None
        out = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_us": k["avg_us"],
               "bytes_per_object": per_obj}
        if name == "pair_tiles":
            # its bytes are far below what HBM allows: the limit is elsewhere
            out["note"] = ("bound by instruction issue and LDS latency in the neighbour scan, not HBM: "
                           "waves wait ~55% of their lifetime, VALU busy about half the kernel "
                           "(profiles/pair_kernel_counters_1M.md). In the loop the pair scans of "
                           "consecutive steps alternate over two streams and overlap; the events "
                           "around a launch also hold the time its workgroups wait for CU slots "
                           "the previous scan still uses (rocprofv3 start-to-end: ~58 us split, "
                           "profiles/kernels_1M_full_model.md; alone: 47.2 us)")
        flops = MEASURED_FLOPS_PER_OBJECT.get((name, form))
        if flops:
            f64 = flops[0] * a.objects / (k["avg_us"] * 1e-6) / 1e12
            f32 = flops[1] * a.objects / (k["avg_us"] * 1e-6) / 1e12
            out["valu"] = {"fp64_tflops": f64, "fp32_tflops": f32, "peak_fp64": FP64_PEAK_TFLOPS,
                           "peak_fp32": FP32_PEAK_TFLOPS,
                           "frac": f64 / FP64_PEAK_TFLOPS + f32 / FP32_PEAK_TFLOPS,
                           "flops_per_object": list(flops)}
        return out

    kern = m["kernels"]
    if kern:  # --profile-steps 0 times no kernel
        dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
        r = roof(main_form, dom, kern[dom])
        if r:
            out["roofline"] = r
    out["roofline_all"] = {form: {name: roof(form, name, k) for name, k in res[form]["kernels"].items()}
                           for form in res}
    out["tuning"] = {form: res[form]["tuning"] for form in res}
    out["kernels_us"] = {form: {k: round(v["avg_us"], 2) for k, v in res[form]["kernels"].items()} for form in res}
    grads = None
    if not a.no_adjoint:
        out["adjoint"], grads = adjoint_leg(la)
        print(f"# adjoint: {json.dumps(out['adjoint'])}", file=sys.stderr, flush=True)
    if not a.no_ensemble:
        etext = (a.ensemble_tuning or "").strip()
        epinned = (json.loads(etext if etext.startswith("{") else Path(etext).read_text()) if etext else None)
        out["ensemble"] = ensemble_leg(la, a.ensemble_steps, pinned=epinned)
        e = out["ensemble"]
        print(f"# ensemble: {e['object_steps_per_s']:.4g} object-steps/s (eager {e['eager_object_steps_per_s']:.4g}),"
              f" graph==eager {e['graph_equals_eager']}, collisions {e['collisions']}", file=sys.stderr, flush=True)
    if not a.no_clustered:
        ctext = (a.clustered_tuning or "").strip()
        cpinned = (json.loads(ctext if ctext.startswith("{") else Path(ctext).read_text()) if ctext else None)
        out["clustered"] = clustered_leg(la, a.clustered_steps, not a.no_cpu_baseline, cpinned)
        print(f"# clustered: {json.dumps(out['clustered'])}", file=sys.stderr, flush=True)
    if dropin is not None:
        out["dropin"] = dropin
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pop, a.cpu_steps)
        if grads is not None:
            out["adjoint"].update(cpu_baseline_adjoint(la, grads))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
