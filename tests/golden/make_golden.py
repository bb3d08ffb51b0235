"""Write the golden fixtures of the README run from a build of the reference.

Run in a container that has the reference sources (it builds oracle/_ref with
oracle/Makefile and runs the harnessed reference main()):

    python tests/golden/make_golden.py

For each run it records, at step N of the time loop (maincode.cpp:461-764):
  * 'pre_va'      the whole state just before TotalVA() of step N
                  (solvent, membrane, parasite, ligands, cells, cell lists);
  * 'post_forces' the forces just after force_calculate_dt() of step N.
Large arrays are stored as SHA-256 digests of their little-endian bytes plus
a strided sample; the 600-vertex membrane, parasite and ligand arrays are
stored whole.  The stdout known answers of the init are kept too.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("IRBC_REFERENCE", "/root/reference")
sys.path.insert(0, HERE)
from refdump import read_dump  # noqa: E402

RUNS = [
    # name, seed, step N, kon, koff, knobs
    ("readme", 12345, 90, "0.005", "0.015", 500),
    ("fastbonds", 777, 40, "40.0", "40.0", 500),
]

WHOLE = {"mx", "my", "mz", "mvx", "mvy", "mvz", "fx", "fy", "fz", "poff", "pon", "ix", "iy",
         "mpid", "bondi", "bondl", "knob", "para", "parai", "glob", "ligi"}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref", "oracle"])
    exe = os.path.join(REPO, "oracle", "_ref", "ref_run")
    meta = {"runs": {}}
    arrays = {}
    for name, seed, n, kon, koff, knobs in RUNS:
        with tempfile.TemporaryDirectory() as tmp:
            for f in ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat"):
                os.symlink(os.path.join(REF, f), os.path.join(tmp, f))
            env = dict(os.environ, OMP_NUM_THREADS="1", IRBC_REF_SEED=str(seed))
            out = subprocess.run([exe, tmp, str(n), kon, koff, "TrophV600.dat", "TrophE600.dat",
                                  "TrophT600.dat", str(knobs), "RBC.dat", "0"], cwd=tmp, env=env,
                                 check=True, capture_output=True, text=True).stdout
            run = {"seed": seed, "step": n, "kon": float(kon), "koff": float(koff),
                   "knobs": knobs, "stdout": out, "hash": {}}
            for phase in ("pre_va", "post_forces"):
                d = read_dump(os.path.join(tmp, "%s_%d.bin" % (phase, n)))
                for k, a in d.items():
                    key = "%s/%s/%s" % (name, phase, k)
                    run["hash"]["%s/%s" % (phase, k)] = digest(a)
                    if k in WHOLE:
                        arrays[key] = a
                    elif a.size > 0:
                        stride = max(1, a.size // 2000)
                        arrays[key + "@sample"] = a[::stride]
                        run["sample_stride_%s_%s" % (phase, k)] = stride
        meta["runs"][name] = run
        print("run", name, "done")
    with open(os.path.join(HERE, "readme_runs.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "readme_runs.npz"), **arrays)
    # the mesh of the README run is a data fixture of the reference
    for f in ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat"):
        with open(os.path.join(REF, f), "rb") as s, open(os.path.join(HERE, f), "wb") as d:
            d.write(s.read())


if __name__ == "__main__":
    main()
