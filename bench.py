"""Benchmark of the iRBC adhesive-dynamics step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-batched]

Headline workload (one GPU): a single replica at the larger size — box
64x64x48 cells at density 10 (1,966,080 MPCD particles, FP32 state with FP64
arithmetic and exact cell sums), a 642-vertex icosphere of radius 8, 500
knobs, kon = 0.005, koff = 0.015, shear rate 0.05 (top wall 2.4).  A step is
one full MD step of the time loop: streaming with bounce-back and walls every
step, an SRD collision every 10th, membrane forces, integrator and bond
kinetics every step, cell->triangle lists every 104th.

Launch options are picked by the library's autotuner on the workload
itself (its trial steps are undone), then the timed run launches without
timing events; a separate pass with HIP events bound to the kernels gives
the per-phase figures and the roofline of the dominant kernel (the
streaming kernel).  Prints ONE JSON line with the throughput, that
roofline, the reference's own CPU path on the host cores for a bounded
sample (with the CPU oracle beside it on the same workload), the batched ensemble (64 replicas of 32x32x24) against
the same replica run alone in replica-steps/s, and the request loop under
Poisson arrivals (p50/p99 latency and throughput).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "irbc-code_amd"))

import ensemble as en  # noqa: E402
import irbc_amd as ia  # noqa: E402
import serve  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step of the FP32 path: streaming reads and
# writes x and v once per step (2 x 24 B); each collision (every csteps = 10
# steps) reads x and v once more for the cell sums (24 B / 10).  The kernel
# roofline charges the streaming launch its 48 B; the step roofline charges
# whole steps 50.4 B.
STREAM_BYTES_PER_PARTICLE = 48.0
STEP_BYTES_PER_PARTICLE = 48.0 + 24.0 / 10.0


def host_cores():
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        ncores = min(ncores, int(os.environ["OMP_NUM_THREADS"]))
    return ncores


REF_RUN = os.path.join(REPO, "oracle", "_ref", "ref_run")
REF_MESH = ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")
REF_PARTICLES = 960000  # 60x40x40 cells at density 10 (define.h:43-48)


def reference_baseline(ncores, budget_s=12.0):
    """The reference's own OpenMP CPU path, its sources compiled by
    oracle/Makefile into oracle/_ref/ref_run, on the host cores: the README
    run (maincode.cpp:56-770), whose box is fixed at compile time
    (define.h:43-45), timed over whole steps of its time loop after its
    set-up, for about budget_s seconds.  None when the build is absent."""
    import subprocess
    import tempfile
    if not os.path.exists(REF_RUN):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        for f in REF_MESH:
            os.symlink(os.path.join(REPO, "tests", "golden", f), os.path.join(tmp, f))
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), IRBC_REF_SEED="12345",
                   IRBC_REF_TIMING=str(budget_s))
        t0 = time.perf_counter()
        r = subprocess.run([REF_RUN, tmp, "100000", "0.005", "0.015", *REF_MESH, "500", "RBC.dat", "0"],
                           cwd=tmp, env=env, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError("reference CPU run failed (%d): %s" % (r.returncode, r.stderr[-500:]))
    tag = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("ref_timing ")]
    if not tag:
        raise RuntimeError("reference CPU run printed no timing line")
    n, sec = int(tag[-1][1]), float(tag[-1][2])
    return {"value": REF_PARTICLES * n / sec, "unit": "particle-steps/s", "cores": int(ncores),
            "kind": "reference",
            "sample": "the reference's OpenMP path (its sources built by oracle/Makefile), README run: "
                      "60x40x40 cells at density 10 (960000 particles, the box define.h fixes), "
                      "Troph600 mesh, 500 knobs, kon 0.005 koff 0.015, seed 12345; %d steps in %.1f s "
                      "after its %.0f s set-up" % (n, sec, wall - sec),
            "ms_per_step": sec * 1e3 / n}


def cpu_baseline(b, d, st, budget_s=10.0, max_steps=5000):
    """The reference's CPU path on the host cores when its build is present,
    with the FP64 oracle beside it on the benchmark's own state for a
    bounded sample; the oracle alone otherwise."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    ncores = host_cores()
    po.set_threads(ncores)
    p, pp = en.params(d), po.Params()
    for f, _ in ia.Params._fields_:
        setattr(pp, f, getattr(p, f))
    o = po.Oracle(pp)
    o.set_mesh(st["V"], st["E"], st["T"])
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
    del o
    out = {"value": ns * n / dt, "unit": "particle-steps/s", "cores": int(ncores), "kind": "port",
           "sample": "FP64 oracle, same 64x64x48 state, %d MD steps (%d collisions), %.1f s"
                     % (n, n // 10, dt),
           "accuracy": accuracy(b, pp, st, po)}
    out["calibration"] = calibration(out["value"], ncores)
    ref = reference_baseline(ncores)
    if ref is None:
        return out
    ref["port"] = out
    ref["oracle_over_reference_here"] = ref["ms_per_step"] / 1e3 * out["value"] / REF_PARTICLES
    ref["accuracy"] = out.pop("accuracy")
    return ref


def calibration(oracle_rate, ncores):
    """The oracle against the reference's own CPU path (its sources, built
    and timed on the same cores in the build container by
    oracle/calibrate.py on the README run): the ratio at the nearest thread
    count, and the reference's rate it implies here."""
    path = os.path.join(REPO, "oracle", "calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        cal = json.load(f)
    run = min(cal["runs"], key=lambda r: abs(r["threads"] - ncores))
    ratio = run["oracle_over_reference"]
    return {"oracle_over_reference": ratio, "at_threads": run["threads"], "measured_on": cal["cpu"],
            "workload": cal["workload"], "reference_particle_steps_per_s": run["reference_particle_steps_per_s"],
            "implied_reference_rate_here": oracle_rate / ratio}


def accuracy(b, pp, st, po):
    """The oracle as a checker on the benchmark's own state: one SRD
    collision with the oracle's draws injected, then the membrane's area,
    volume and forces; relative errors of the GPU path (FP32-stored solvent,
    exact fixed-point cell sums, FP64 membrane) against the FP64 oracle."""
    o = po.Oracle(pp)
    o.set_mesh(st["V"], st["E"], st["T"])
    s = b.solvent(0)
    o.set_solvent(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
    v = b.vertices(0)
    o.set_vertices(**{k: v[k] for k in po.Oracle.VD + po.Oracle.VI})
    o.set_parasite(*b.parasite(0))
    o.set_ligands(b.ligands(0))
    w = b.globals(0)["wallspeed"]
    o.set_wallspeed(w)
    o.seed_mt(11, 22)
    o.collision()
    b.set_diagnostics(True)
    b.set_collision_tape(0, *o.get_collision_tape())
    b.collision_step()
    co, cg = o.cells(), b.cells(0)
    po_ = co["vcm"].reshape(-1, 3) * co["ncc"][:, None]
    pg_ = cg["vcm"].reshape(-1, 3) * cg["ncc"][:, None]
    cell_p = float(np.abs(pg_ - po_).max() / np.abs(po_).max())
    so, sg = o.solvent(), b.solvent(0)
    vel = max(float(np.abs(sg[k] - so[k]).max() / np.abs(so[k]).max()) for k in ("vx", "vy", "vz"))
    keys_equal = bool(np.array_equal(sg["key"], so["pid"]) and np.array_equal(cg["nc"], co["nc"]))
    o.totalva()
    o.pfdt()
    o.forces()
    o.external()
    b.membrane_forces()
    go, gg = o.globals(), b.globals(0)
    vo, vg = o.vertices(), b.vertices(0)
    force = max(float(np.abs(vg[k] - vo[k]).max() / np.abs(vo[k]).max()) for k in ("fx", "fy", "fz"))
    return {"cell_momentum_rel": cell_p, "rotated_velocity_rel_fp32_storage": vel,
            "cell_keys_and_counts_bit_exact": keys_equal,
            "area_rel": abs(gg["Sarea"] - go["Sarea"]) / go["Sarea"],
            "volume_rel": abs(gg["Volume"] - go["Volume"]) / go["Volume"],
            "membrane_force_rel": force,
            "targets": {"cell_momentum_rel": 1e-5, "area_volume_rel": 1e-8}}


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


def timed(b, steps, warmup):
    b.md_steps(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.md_steps(steps)
    b.sync()
    return time.perf_counter() - t0


def set_options(b, k):
    for i, (_, opt) in enumerate(ia.OPTIONS):
        b.set_option(opt, (k >> i) & 1)
    return k


def batched(steps, warmup, options=-1):
    """64 replicas of 32x32x24 in one batch vs the first of them alone
    (each autotuned, or both with the given launch-option index)."""
    ds = en.uniform()
    b = en.batch(ds)
    kb = set_options(b, options) if options >= 0 else b.autotune(20)[0]
    span_b = b.option(ia.OPT_BIN_SPAN)
    dt_b = timed(b, steps, warmup)
    np_b = sum(en.n_solvent(d) for d in ds)
    b.close()
    a = en.batch(ds[:1])
    ka = set_options(a, options) if options >= 0 else a.autotune(50)[0]
    dt_a = timed(a, steps, warmup)
    a.close()
    rb = len(ds) * steps / dt_b
    ra = steps / dt_a
    return {"workload": "64 replicas of 32x32x24 cells, density 10 (245760 particles each), "
                        "642-vertex icosphere r=6, 200-800 knobs, kon/koff log-uniform [1e-3,1e-1]",
            "steps": steps, "warmup": warmup,
            "batched_replica_steps_per_s": rb, "plain_replica_steps_per_s": ra,
            "speedup": rb / ra, "batched_ms_per_step": dt_b * 1e3 / steps,
            "plain_ms_per_step": dt_a * 1e3 / steps,
            "batched_particle_steps_per_s": np_b * steps / dt_b,
            "launch_options_index": {"batched": kb, "plain": ka}, "bin_span": {"batched": span_b}}


def plain_rate(ds, steps, warmup=50):
    """Replica-steps/s of the plain path (one replica per context, the
    library's default launch shape) over the replicas ds, run one after
    another: total steps over total time."""
    t, n = 0.0, 0
    for d in ds:
        a = en.batch([d])
        a.md_steps(warmup)
        a.sync()
        t0 = time.perf_counter()
        a.md_steps(steps)
        a.sync()
        t += time.perf_counter() - t0
        n += steps
        a.close()
    return n / t


def served(n, rate, steps, budget_gb, plain_every=8, plain_steps=1000, options=None, repack=True, chunk=250,
           window=0.1, repack_idle=0.25, bin_span=1):
    """Poisson arrivals of uneven-ensemble requests through the request
    loop, against the plain path on every plain_every-th request (a sample
    of the same mix, run one after another); contexts of several replicas
    take the launch-option index `options` (None: the library's defaults).
    repack: every running request in one context, re-packed at admissions
    and finishes; else micro-batches side by side."""
    reqs = serve.poisson_requests(n, rate, steps, seed=1)
    srv = serve.Server(int(budget_gb * 2 ** 30), chunk=chunk, max_active=2, window=window, options=options,
                       repack=repack, repack_idle=repack_idle, bin_span=bin_span)
    srv.run(reqs)
    out = serve.summary(reqs, srv.makespan)
    sample = [r.replica for r in reqs[::plain_every]]
    out.update({"workload": "%d requests, Poisson arrivals at %.1f/s, %d steps each, boxes 24^3..64x64x48 "
                            "at density 5-10, 162/642/2562-vertex ellipsoids" % (n, rate, steps),
                "schedule": "re-packing: one running context, rebuilt at admissions and finishes" if repack
                else "micro-batches side by side",
                "hbm_budget_GiB": budget_gb, "chunk_steps": srv.chunk, "batching_window_s": srv.window,
                "repack_idle_share": srv.repack_idle if repack else None,
                "contexts_built": srv.stats["groups"], "max_concurrent_contexts": srv.stats["max_active_groups"],
                "max_resident_replicas": srv.stats["max_resident_replicas"],
                "launch_options_index": options, "bin_span": bin_span,
                "plain_replica_steps_per_s": plain_rate(sample, plain_steps),
                "plain_sample": "every %dth request alone (%d of them), %d timed steps each"
                                % (plain_every, len(sample), plain_steps)})
    out["speedup_vs_plain"] = out["replica_steps_per_s"] / out["plain_replica_steps_per_s"]
    if repack:
        out["idle_waiting_for_build_s"] = srv.idle
        out["contexts"] = srv.log
        out["chunks"] = srv.chunks
    return out


def uneven(steps, plain_every=6, plain_steps=1000, options=-1):
    """48 replicas of mixed boxes, meshes, rates and shears in one batch
    (list updates at collision steps), autotuned (or with the launch-option
    index `options`), against the plain path on every plain_every-th of
    them (plain_every 0: no plain sample, for profiling the batch alone)."""
    ds = en.uneven()
    b = en.batch(ds)
    b.set_option(ia.OPT_ALIGN_LISTS, 1)
    kb = set_options(b, options) if options >= 0 else b.autotune(20)[0]
    span, span_ms = b.option(ia.OPT_BIN_SPAN), _span_ms(b)
    dt = timed(b, steps, 50)
    b.close()
    ns = sum(en.n_solvent(d) for d in ds)
    rb = len(ds) * steps / dt
    if plain_every <= 0:
        return {"steps": steps, "batched_ms_per_step": dt * 1e3 / steps, "batched_replica_steps_per_s": rb,
                "batched_particle_steps_per_s": ns * steps / dt, "launch_options_index": kb, "bin_span": span}
    ra = plain_rate(ds[::plain_every], plain_steps)
    return {"workload": "48 replicas, boxes 24^3..64x64x48 at density 5-10 (%d particles), 162/642/2562-vertex "
                        "ellipsoids, 100-2000 knobs, shear 0.01-0.1" % ns,
            "steps": steps, "batched_ms_per_step": dt * 1e3 / steps, "batched_replica_steps_per_s": rb,
            "batched_particle_steps_per_s": ns * steps / dt, "plain_replica_steps_per_s": ra,
            "plain_sample": "every %dth replica alone (%d of them), %d timed steps each"
                            % (plain_every, len(ds[::plain_every]), plain_steps),
            "speedup": rb / ra, "launch_options_index": kb, "bin_span": span, "bin_span_ms_per_step": span_ms}


def _span_ms(b):
    """the autotuner's ms per step with bin spans 1, 2, 4 (None: not timed)"""
    return dict((k, round(float(v), 4) if v == v else None) for k, v in zip((1, 2, 4), b.autotune_spans()))


def trace_summary(path, skip, steps):
    """Timeline of `steps` MD steps from a rocprofv3 kernel trace (CSV),
    after the first `skip` fused streaming launches: every kernel with its
    start, duration and the idle gap before it, then the span, the busy
    time and their per-step averages.  Returns the text."""
    import csv
    import re
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    ks = [(re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", ""), int(r["Start_Timestamp"]),
           int(r["End_Timestamp"])) for r in rows]
    lead = ("k_step_stream", "k_stream")
    idx = [i for i, k in enumerate(ks) if k[0].startswith(lead)]
    if len(idx) < skip + steps:
        raise SystemExit("trace holds %d streaming launches, fewer than %d" % (len(idx), skip + steps))
    # from the first kernel after the previous step's last launch (a call's
    # prologue belongs to the step it precedes) to the last step's end
    j = idx[skip]
    while j > 0 and not ks[j - 1][0].startswith(("k_step_near", "k_near")):
        j -= 1
    e = idx[skip + steps] if skip + steps < len(idx) else len(ks)
    while e > 0 and not ks[e - 1][0].startswith(("k_step_near", "k_near", "k_mem")):
        e -= 1
    t0, prev, busy, by = ks[j][1], ks[j][1], 0, {}
    lines = ["%-24s %10s %9s %8s" % ("kernel", "start_us", "dur_us", "gap_us")]
    for name, a, b in ks[j:e]:
        lines.append("%-24s %10.1f %9.1f %8.1f" % (name[:24], (a - t0) / 1e3, (b - a) / 1e3, max(0, a - prev) / 1e3))
        busy += b - a
        by[name] = by.get(name, 0) + (b - a)
        prev = max(prev, b)
    span = prev - t0
    lines.append("span %.1f us, busy %.1f us, idle %.1f us over %d steps: %.2f us per step "
                 "(busy %.2f, idle %.2f)" % (span / 1e3, busy / 1e3, (span - busy) / 1e3, steps,
                                             span / 1e3 / steps, busy / 1e3 / steps, (span - busy) / 1e3 / steps))
    lines.append("busy time per kernel, us per step:")
    for name, t in sorted(by.items(), key=lambda kv: -kv[1]):
        lines.append("  %-24s %8.2f" % (name[:24], t / 1e3 / steps))
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--uneven-steps", type=int, default=1000)
    ap.add_argument("--uneven-plain-every", type=int, default=6,
                    help="time every n-th replica of the uneven batch alone (0: none)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the 64-replica batch against its first replica alone (for profiling)")
    ap.add_argument("--only", choices=("batched", "uneven", "served"),
                    help="run only that leg and print its JSON object (for profiling)")
    ap.add_argument("--served-requests", type=int, default=64)
    ap.add_argument("--served-rate", type=float, default=32.0)
    ap.add_argument("--served-steps", type=int, default=5000)
    ap.add_argument("--served-chunk", type=int, default=250)
    ap.add_argument("--served-window", type=float, default=0.1)
    ap.add_argument("--served-budget-gb", type=float, default=64.0)
    ap.add_argument("--served-span", type=int, default=1, choices=(1, 2, 4),
                    help="bin-sort tile span of contexts of several replicas (the full run takes the uneven leg's)")
    ap.add_argument("--served-idle", type=float, default=0.25,
                    help="re-pack once finished requests hold this share of the running context's particles")
    ap.add_argument("--served-microbatches", action="store_true",
                    help="serve with micro-batches side by side instead of re-packing")
    ap.add_argument("--profile-every", type=int, default=4)
    ap.add_argument("--options", type=int, default=-1,
                    help="launch-option index (bit i = irbc_amd.OPTIONS[i]) instead of autotuning")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="summarise a rocprofv3 kernel trace of this benchmark (the --steps steps after "
                         "--warmup steps) and exit; needs no GPU")
    a = ap.parse_args()
    if a.trace_summary:
        print(trace_summary(a.trace_summary, a.warmup, a.steps))
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the replicas of this workload are independent "
              "and shard by replica, so there is no multi-GPU run to measure (asked for %d)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    ia.load()
    if a.batched_only or a.only == "batched":
        print(json.dumps({"batched": batched(a.batched_steps, 100, a.options)}))
        return
    if a.only == "uneven":
        print(json.dumps({"uneven": uneven(a.uneven_steps, plain_every=a.uneven_plain_every, options=a.options)}))
        return
    if a.only == "served":
        print(json.dumps({"served": served(a.served_requests, a.served_rate, a.served_steps, a.served_budget_gb,
                                           options=None if a.options < 0 else a.options,
                                           repack=not a.served_microbatches, chunk=a.served_chunk,
                                           window=a.served_window, repack_idle=a.served_idle,
                                           bin_span=a.served_span)}))
        return
    d = en.single()
    b = ia.Batch([en.spec(d)])
    st = en.load(b, 0, d)
    en.start(b)
    ns = en.n_solvent(d)
    if a.options >= 0:
        tuned, trials = a.options, np.zeros(0)
        for i, (_, opt) in enumerate(ia.OPTIONS):
            b.set_option(opt, (tuned >> i) & 1)
    else:
        tuned, trials = b.autotune(50)  # state is restored: the trajectory does not advance
    span, span_ms = b.option(ia.OPT_BIN_SPAN), _span_ms(b)
    b.md_steps(a.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.md_steps(a.steps)
    b.sync()
    dt = time.perf_counter() - t0
    # a separate, shorter pass with HIP events on the kernels (launch by
    # launch) for the per-phase figures and the roofline of k_stream
    b.profile(a.profile_every)
    b.md_steps(min(a.steps, 4000))
    b.sync()
    prof = b.profile_read()
    b.profile(0)
    value = ns * a.steps / dt
    k_ms = prof["stream_ms"] / max(prof["stream_n"], 1)
    achieved = STREAM_BYTES_PER_PARTICLE * ns / (k_ms * 1e-3) / 1e9
    step_gbs = STEP_BYTES_PER_PARTICLE * value / 1e9
    opt = dict((name, (tuned >> i) & 1) for i, (name, _) in enumerate(ia.OPTIONS))
    if opt["one_queue"] and opt["membrane_chain"]:
        kname = "k_step_stream<%s>" % ("true" if opt["stream4"] else "false")
    else:
        kname = "k_stream4" if opt["stream4"] else "k_stream<false>"
    tfile = "%s_c2_traffic.md" % kname.replace("<", "_").replace(">", "")
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
                     "traffic": load_traffic(tfile), "bytes_per_particle": STREAM_BYTES_PER_PARTICLE,
                     "kernel": kname, "kernel_ms": k_ms, "kernel_launches_timed": prof["stream_n"],
                     "note": "kernel_ms is the whole fused launch: the bulk sweep and, on one queue, "
                             "the membrane chain's phases riding in it",
                     "step": {"achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                              "bytes_per_particle_step": STEP_BYTES_PER_PARTICLE,
                              "ms_per_step": dt * 1e3 / a.steps}},
        "launch_options": dict([(name, (tuned >> i) & 1) for i, (name, _) in enumerate(ia.OPTIONS)],
                               bin_span=span, bin_span_ms_per_step=span_ms,
                               autotune_ms_per_step=[round(x, 4) if x == x else None for x in trials.tolist()]),
        "phases_ms_per_launch": {
            "stream": k_ms,
            "collision_sequence": prof["collision_ms"] / prof["collision_n"] if prof["collision_n"] else None,
            # on one queue the membrane runs inside the streaming launch
            "membrane": prof["membrane_ms"] / prof["membrane_n"] if prof["membrane_n"] else None,
            "stream_near": prof["near_ms"] / prof["near_n"] if prof["near_n"] else None},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(b, d, st)
    b.close()
    if not a.no_batched:
        line["batched"] = batched(a.batched_steps, 100)
    if not a.no_batched:
        line["uneven"] = uneven(a.uneven_steps)
    if not a.no_served:
        # micro-batches of several replicas take the launch shape the
        # autotuner picked for the uneven ensemble, the same mix
        opt = line["uneven"]["launch_options_index"] if "uneven" in line else None
        a.served_span = line["uneven"]["bin_span"] if "uneven" in line else 1
        line["served"] = served(a.served_requests, a.served_rate, a.served_steps, a.served_budget_gb, options=opt,
                                repack=not a.served_microbatches, chunk=a.served_chunk, window=a.served_window,
                                repack_idle=a.served_idle, bin_span=a.served_span)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
