"""Benchmark of the sparse fire step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fire256_8m]

A step is one HNS_Advance call (one substep) of the fire scene: advect velocity, vorticity
confinement, divergence + combustion + buoyancy, 100 red-black Gauss-Seidel iterations,
gradient subtraction, MacCormack advection of the five scalars. The inputs are resident in HBM
before timing starts (the context holds the state between calls). Prints one JSON line.
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

METRIC = "voxel-steps/s at 8M voxels to 1e-5 residual, plain vs learned; % HBM peak"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one RBGS colour sweep per voxel of the domain: read the other colour's p
# (2 B), own div (2 B), own p (2 B), write own p (2 B)
RBGS_SWEEP_BYTES = 8.0


def step_bytes(K: int, m_scalars: int) -> float:
    """Compulsory bytes per voxel-step of the fire step (FP32; stencil reuse perfect):
    adv u 24 + vorticity 24 + div/combustion/buoyancy 60 + 16 K + gradient 28 + scalars (12 + 8 M)."""
    return 24 + 24 + 60 + 16 * K + 28 + 12 + 8 * m_scalars


def make_scene(name):
    from hnanosolver_amd import scenes
    return scenes.SCENES[name]()


def payload(sc):
    import hnanosolver_amd as hns
    d = hns.GridIndexedData(sc.coords)
    for k, v in sc.fields.items():
        d.add_value_block(k, v)
    d.add_value_block("velocity", sc.velocity)
    return d


def cpu_baseline(sc, budget_s: float = 20.0):
    """The CPU oracle (a from-scratch restatement of the reference step; the reference has no CPU
    solver) on the host cores, on a bounded sample: whole steps of the same workload until about
    budget_s seconds have run."""
    import oracle
    o = oracle.Oracle(sc.coords)
    vel, f = sc.velocity, dict(sc.fields)
    t0 = time.perf_counter()
    steps = 0
    while True:
        vel, f = o.advance(sc.voxel_size, sc.iterations, sc.dt, sc.params, vel, f)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 3:
            break
    return {"value": sc.n * steps / el, "unit": "voxel-steps/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": "%d full step(s) of %s (N=%d, K=%d) on the CPU oracle in %.1f s" % (steps, sc.name, sc.n,
                                                                                     sc.iterations, el)}


def read_traffic(workload):
    path = os.path.join(ROOT, "profiles", "rbgs_%s_traffic.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="fire256_8m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the step runs on a single MI355X (sharding across GPUs is a "
              "design in DESIGN.md, not built); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import hnanosolver_amd as hns

    print("[bench] building scene %s" % a.workload, file=sys.stderr, flush=True)
    sc = make_scene(a.workload)
    data = payload(sc)
    topo = hns.create_index_grid(data, sc.voxel_size)
    ctx = hns.ensure_context(None, data, topo, sc.voxel_size)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    params = hns.make_params(**sc.params)
    K = sc.iterations
    print("[bench] N=%d leaves=%d K=%d; warmup %d" % (sc.n, sc.n_leaves, K, a.warmup), file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        hns.advance(ctx, K, sc.dt, params)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        hns.advance(ctx, K, sc.dt, params)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / a.steps
    value = sc.n / (ms * 1e-3)
    print("[bench] timed %d steps: %.3f ms/step (wall %.3f ms/step)" % (a.steps, ms, wall * 1e3 / a.steps),
          file=sys.stderr, flush=True)

    # instrumented pass: HIP events around every launch of each kernel class, same step count
    ctx.set_timing(True)
    ctx.reset_timing()
    for _ in range(a.steps):
        hns.advance(ctx, K, sc.dt, params)
    torch.cuda.synchronize()
    tm = ctx.timing()
    ctx.set_timing(False)
    rb_ms, rb_n = tm["rbgs"]
    avg_s = rb_ms * 1e-3 / max(rb_n, 1)
    algo = RBGS_SWEEP_BYTES * sc.n
    achieved = algo / avg_s / 1e9
    kernels = {k: {"ms_per_step": v[0] / a.steps, "launches_per_step": v[1] / a.steps,
                   "avg_us": 1e3 * v[0] / max(v[1], 1)} for k, v in tm.items() if v[1]}
    m_scalars = sum(1 for k in ("density", "temperature", "fuel", "waste", "flame") if ctx.has_field(k))
    step_b = step_bytes(K, m_scalars)
    out = {
        "metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": sc.name, "voxels": sc.n, "leaves": sc.n_leaves, "rbgs_iterations": K, "dt": sc.dt,
                   "voxel_size": sc.voxel_size, "advected_scalars": m_scalars, "solver": "rbgs_fixed",
                   "notes": sc.notes},
        "roofline": {"bound": "hbm", "kernel": "rbgs_sweep", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(sc.name),
                     "algorithmic_bytes_per_launch": algo, "avg_launch_us": avg_s * 1e6},
        "step_roofline": {"bytes_per_voxel_step": step_b, "achieved_GBs": step_b * value / 1e9,
                          "frac": step_b * value / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
    }
    hns.destroy_context(ctx)
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, a.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
