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
    p.add_argument("--forms", default="fp64,split")
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
    cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 24)
    return la.Engine(cfg)


def run_form(la, pop, precision, steps, warmup, profile_steps):
    eng = make_engine(la, precision)
    eng.set_objects(pop)
    # plane-tile shape picked by timing the scan on the initial state (outside the timed region)
    tuning = eng.autotune(repetitions=3)
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
    n = pop.n
    r, v = pop.r(), pop.v()
    cfg = ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0)
    t0 = time.perf_counter()
    for s in range(steps):
        r, v = ol.propagate(r, v, s * 10.0, 10.0, 0xFF, 1, pop.aom, pop.bc_inv, pop.cr, pop.activity, threads)
        ol.pairs(ol.ObjectSet(r, v, pop.radius_m, pop.activity, pop.id, pop.parent), cfg, "grid", threads)
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "object-steps/s", "cores": threads, "kind": "port",
            "sample": f"C oracle (OpenMP, -O2, no FMA contraction): {n} objects x {steps} steps of the same "
                      f"workload (propagate + linked-cell scan), {el:.1f} s on {threads} host threads"}


def adjoint_leg(la, reps=3):
    """Differentiable propagation: 4,096 objects, full force model in FP64, dt = 60 s, 1,440
    steps (one day), checkpoints every 64 steps; loss = sum |r_T - target|^2 with targets =
    nominal final state + N(0, 100 m).  Reports gradient-steps/s (objects x steps per
    forward + backward pass, device time from events on torch's stream) and the relative
    error of the gradient against the oracle's FP64 complex-step gradient."""
    import torch
    from ladds_amd import adjoint

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol

    n, steps, dt = 4096, 1440, 60.0
    pop = la.generate_population(1, 0x1ADD5 + 4, n)
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    plan = adjoint.AdjointPlan(n, steps, dt, components=la.ALL_COMPONENTS, checkpoint_every=64)
    with torch.no_grad():
        rT0, _ = adjoint.propagate(t(pop.r()), t(pop.v()), t(pop.bc_inv), plan, aom=t(pop.aom), cr=t(pop.cr))
    target = rT0 + 0.1 * torch.randn(rT0.shape, dtype=torch.float64, device=dev,
                                     generator=torch.Generator(device=dev).manual_seed(4))
    times = []
    aom_t, cr_t = t(pop.aom), t(pop.cr)
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
    g_r = 2.0 * (rT.detach().cpu().numpy() - target.cpu().numpy())
    t0 = time.perf_counter()
    threads = min(16, len(os.sched_getaffinity(0)))
    cr_, cv_, cbc = ol.vjp(pop.r(), pop.v(), 0.0, dt, la.ALL_COMPONENTS, steps, g_r, np.zeros_like(g_r),
                           aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr, nthreads=threads)
    cpu_sec = time.perf_counter() - t0
    got = np.concatenate([r0.grad.cpu().numpy(), v0.grad.cpu().numpy()], 1)
    ref = np.concatenate([cr_, cv_], 1)
    rel = np.max(np.abs(got - ref), 1) / np.max(np.abs(ref), 1)
    bc_rel = np.abs(bc.grad.cpu().numpy() - cbc) / np.max(np.abs(cbc))
    plan.close()
    return {"workload": f"{n} objects, full force model FP64, dt {dt:g} s, {steps} steps, checkpoints every 64",
            "gradient_steps_per_s": n * steps / sec, "ms_per_gradient": sec * 1e3,
            "grad_rel_err_max": float(rel.max()), "grad_rel_err_median": float(np.median(rel)),
            "bc_grad_rel_err_max": float(bc_rel.max()),
            "cpu_reference": {"kind": "port", "method": "complex-step FP64 (7 complex propagations per object)",
                              "gradient_steps_per_s": n * steps / cpu_sec, "cores": threads}}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the engine has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
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
    for name in a.forms.split(","):
        res[name] = run_form(la, pop, forms[name], a.steps, a.warmup, a.profile_steps)
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
    # roofline of the dominant kernel of the headline form
    kern = m["kernels"]
    dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    per_obj = BYTES.get(dom, {}).get(main_form)
    if per_obj:
        achieved = per_obj * a.objects / (kern[dom]["avg_us"] * 1e-6) / 1e9
        out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "avg_us": kern[dom]["avg_us"], "bytes_per_object": per_obj}
    out["tuning"] = {form: res[form]["tuning"] for form in res}
    out["kernels_us"] = {form: {k: round(v["avg_us"], 2) for k, v in res[form]["kernels"].items()} for form in res}
    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(la)
        print(f"# adjoint: {json.dumps(out['adjoint'])}", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pop, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
