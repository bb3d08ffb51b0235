"""Benchmark of the iRBC adhesive-dynamics step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config c2|c3|c3plain] [--no-cpu]

Default workload (one GPU): a single replica at the larger size — box
64x64x48 cells at density 10 (1,966,080 MPCD particles, FP32 state with FP64
arithmetic and exact cell sums), a 642-vertex icosphere of radius 8, 500
knobs, kon = 0.005, koff = 0.015, shear rate 0.05 (top wall 2.4).  A step is
one full MD step of the time loop: streaming with bounce-back and walls every
step, an SRD collision every 10th, membrane forces, integrator and bond
kinetics every step, cell->triangle lists every 104th.

Prints ONE JSON line with the throughput, the HBM roofline of the dominant
kernel (the streaming kernel, timed with HIP events on its own stream) and
the CPU oracle's rate on the host cores for a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "irbc-code_amd"))

import irbc_amd as ia  # noqa: E402
import scenario as sc  # noqa: E402

HBM_PEAK_GBS = 8000.0
STREAM_BYTES_PER_PARTICLE = 48.0  # x,y,z,vx,vy,vz FP32 read + write


def c2_params():
    return dict(Lx=64, Ly=64, Lz=48, walls=0.05 * 48, ramp_steps=0, onrate=0.005, offrate=0.015)


def build_replica(b, r, kw, nv, radius, nknobs, seed, density=10):
    p = ia.Params.default(**kw)
    V, E, T = sc.mesh_for(nv)
    V = sc.place_reference(V * radius, p.Lx, p.Ly)
    ns = int(density * p.Lx * p.Ly * p.Lz)
    sol = sc.solvent(ns, p.Lx, p.Ly, p.Lz, seed)
    vel = sc.vertex_velocities(nv, seed + 1, pmass=p.pmass)
    knob = sc.knobs_uniform(nv, nknobs, seed + 2)
    pd, pi = sc.parasite_state(V, seed + 3)
    b.set_mesh(r, V, E, T)
    b.set_solvent(r, *sol)
    b.set_vertices(r, vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2], knob=knob)
    b.set_parasite(r, pd, pi)
    b.set_counters(r, 0, -1, 0)
    return dict(V=V, E=E, T=T, sol=sol, vel=vel, knob=knob, para=(pd, pi), params=kw)


def spec_for(kw, nv, seed, density=10):
    p = ia.Params.default(**kw)
    ns = int(density * p.Lx * p.Ly * p.Lz)
    V, E, T = sc.mesh_for(nv)
    return ia.make_spec(p, ns, nv, len(E), len(T), seed=seed)


def init_batch(b):
    """maincode.cpp:383-391: forces at rest, inside/outside flags, cell lists."""
    b.membrane_phases(ia.PH_VA | ia.PH_FORCES)
    b.classify()
    b.llist_update()
    b.sync()


def cpu_baseline(b, setup, budget_s=15.0, max_steps=5000):
    """The FP64 oracle on the host cores, on the same state, for a bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        ncores = min(ncores, int(os.environ["OMP_NUM_THREADS"]))
    po.set_threads(ncores)
    o = po.Oracle(po.Params.default(**setup["params"]))
    o.set_mesh(setup["V"], setup["E"], setup["T"])
    s = b.solvent(0)
    o.set_solvent(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
    v = b.vertices(0)
    o.set_vertices(**{k: v[k] for k in po.Oracle.VD + po.Oracle.VI})
    o.set_parasite(*b.parasite(0))
    o.llistupdate()
    o.set_counters(0, -1, 0)
    o.md_step()  # first-touch warm-up
    n, t0 = 0, time.perf_counter()
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        o.md_step()
        n += 1
    dt = time.perf_counter() - t0
    ns = len(s["x"])
    return {"value": ns * n / dt, "unit": "particle-steps/s", "cores": int(ncores), "kind": "port",
            "sample": "FP64 oracle, same 64x64x48 state, %d MD steps (%d collisions), %.1f s"
                      % (n, (n + 1) // 10, dt)}


def load_traffic(name):
    """HBM bytes per launch from a committed counter summary (its
    `bytes_per_launch:` line), or None when there is none."""
    path = os.path.join(REPO, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-every", type=int, default=4)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the replicas of this workload are independent "
              "and shard by replica, so there is no multi-GPU run to measure (asked for %d)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    ia.load()
    kw = c2_params()
    spec = spec_for(kw, 642, 12345)
    b = ia.Batch([spec])
    setup = build_replica(b, 0, kw, 642, 8.0, 500, 12345)
    init_batch(b)
    ns = spec.n_solvent
    b.md_steps(a.warmup)
    b.sync()
    b.profile(a.profile_every)
    t0 = time.perf_counter()
    b.md_steps(a.steps)
    b.sync()
    dt = time.perf_counter() - t0
    prof = b.profile_read()
    b.profile(0)
    value = ns * a.steps / dt
    k_ms = prof["stream_ms"] / max(prof["stream_n"], 1)
    achieved = STREAM_BYTES_PER_PARTICLE * ns / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "MPCD particle-steps/s", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single replica 64x64x48 cells, density 10 (1966080 particles), "
                               "642-vertex icosphere r=8, 500 knobs, kon 0.005 koff 0.015, "
                               "shear rate 0.05", "storage": "fp32 solvent SoA, fp64 membrane",
                   "replicas": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic("k_stream_c2_traffic.md"),
                     "kernel": "k_stream", "kernel_ms": k_ms, "kernel_launches_timed": prof["stream_n"]},
        "phases_ms_per_launch": {
            "stream": k_ms,
            "collision_sequence": prof["collision_ms"] / max(prof["collision_n"], 1),
            "membrane": prof["membrane_ms"] / max(prof["membrane_n"], 1),
            "stream_near": prof["near_ms"] / max(prof["near_n"], 1)},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(b, setup)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
