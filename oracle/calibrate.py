"""Time the reference's CPU path and the oracle on the same cores.

TEST INFRASTRUCTURE ONLY.  Run in a container that has the reference sources
(it builds oracle/_ref/ref_run, the reference's own sources compiled by
oracle/Makefile, and runs the harnessed reference main()):

    python oracle/calibrate.py [--threads 8 1] [--steps 10 110]

Workload: the README run (maincode.cpp:56-770): 60x40x40 at density 10
(960 000 particles), the Troph600 mesh, 500 knobs, kon 0.005, koff 0.015,
seed 12345, wall speed 0.1 (every step of the first 250 000).  The
reference's time per step is taken between the end of its set-up (about
46 s, the serial inside/outside test; knobs.dat is written last) and the
state dump of step n, in two runs that stop after n0 and n1 steps, so the
dump's own time cancels; the oracle's steps are timed directly, after the
same set-up.  The result, with the ratio
oracle/reference per thread count, goes to oracle/calibration.json, which
the benchmark quotes beside the oracle's own timing on the GPU box's cores.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF = os.environ.get("IRBC_REFERENCE", "/root/reference")
GOLD = os.path.join(REPO, "tests", "golden")
MESH = ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")
NS = 960000


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def ref_run_seconds(exe, n, threads):
    """Seconds from the end of the reference's set-up (knobs.dat written,
    maincode.cpp:436-446) to the end of the state dump at step n."""
    with tempfile.TemporaryDirectory() as tmp:
        for f in MESH:
            os.symlink(os.path.join(GOLD, f), os.path.join(tmp, f))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), IRBC_REF_SEED="12345")
        subprocess.run([exe, tmp, str(n), "0.005", "0.015", *MESH, "500", "RBC.dat", "0"], cwd=tmp, env=env,
                       check=True, capture_output=True)
        return (os.stat(os.path.join(tmp, "pre_va_%d.bin" % n)).st_mtime_ns
                - os.stat(os.path.join(tmp, "knobs.dat")).st_mtime_ns) * 1e-9


def oracle_step_seconds(n, threads):
    sys.path.insert(0, HERE)
    import pyoracle as po
    po.set_threads(threads)
    V, E, T = po.read_mesh(*[os.path.join(GOLD, f) for f in MESH])
    o = po.Oracle(onrate=0.005, offrate=0.015)
    o.init_readme(12345, NS, V, E, T, 500)
    t0 = time.perf_counter()
    for _ in range(n):
        o.md_step()
    return (time.perf_counter() - t0) / n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, nargs="+", default=[8, 1])
    ap.add_argument("--steps", type=int, nargs=2, default=[10, 110])
    a = ap.parse_args()
    subprocess.check_call(["make", "-s", "-C", HERE, "ref", "oracle"])
    exe = os.path.join(HERE, "_ref", "ref_run")
    n0, n1 = a.steps
    out = {"workload": "README run 60x40x40, 960000 particles, Troph600 mesh, 500 knobs, kon 0.005 "
                       "koff 0.015, seed 12345, wall speed 0.1",
           "cpu": cpu_model(), "nproc": os.cpu_count(), "runs": []}
    for k in a.threads:
        t0 = ref_run_seconds(exe, n0, k)
        t1 = ref_run_seconds(exe, n1, k)
        ref = (t1 - t0) / (n1 - n0)
        orc = oracle_step_seconds(n1 - n0, k)
        row = {"threads": k, "reference_ms_per_step": ref * 1e3, "oracle_ms_per_step": orc * 1e3,
               "reference_particle_steps_per_s": NS / ref, "oracle_particle_steps_per_s": NS / orc,
               "oracle_over_reference": ref / orc, "steps": n1 - n0}
        print(json.dumps(row), flush=True)
        out["runs"].append(row)
    with open(os.path.join(HERE, "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
