"""Build the MI355X library, the host driver and the CPU oracle, in-tree.

    python irbc-code_amd/build.py

Outputs:
  irbc-code_amd/libirbc_amd.so   HIP kernels + C ABI (hipcc, gfx950)
  irbc-code_amd/libirbc_host.so  scenario set-up, reference-named functions
  irbc-code_amd/irbc_main        reference-style command-line driver
  irbc-code_amd/copy_probe       HBM copy-bandwidth probe (counter calibration)
  oracle/liboracle.so            CPU checker (test infrastructure)
  oracle/_ref/ref_run            harnessed reference build (only where the
                                 reference sources are present)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: no FMA contraction, so FP64 arithmetic rounds like the
# reference's expression by expression; packed FP32 ops disabled on purpose.
HIPFLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_lib():
    out = os.path.join(HERE, "libirbc_amd.so")
    tmp = out + ".tmp"
    # the soname lets a process that loaded this file by path and then
    # libirbc_host.so (which needs it by name) share one copy
    run([HIPCC] + HIPFLAGS + ["-shared", "-Wl,-soname,libirbc_amd.so", "-o", tmp,
                              os.path.join(HERE, "csrc", "irbc.hip")])
    os.replace(tmp, out)
    return out


def build_probe():
    """Copy-bandwidth probe used to calibrate the HBM traffic counters."""
    run([HIPCC] + HIPFLAGS + ["-o", os.path.join(HERE, "copy_probe"),
                              os.path.join(HERE, "csrc", "copy_probe.hip")])


def build_host():
    """Host library (scenario set-up, reference-named functions, C entry
    points) and the command-line driver, both over libirbc_amd.so."""
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + os.path.join(HERE, "host")]
    lib = ["-L" + HERE, "-lirbc_amd", "-Wl,-rpath,$ORIGIN"]
    src = [os.path.join(HERE, "host", f) for f in ("scenario.cpp", "refnames.cpp", "host_capi.cpp")]
    out = os.path.join(HERE, "libirbc_host.so")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc + ["-o", out + ".tmp"] + src + lib)
    os.replace(out + ".tmp", out)
    run(["g++", "-O2", "-std=c++17", "-Wall"] + inc + ["-o", os.path.join(HERE, "irbc_main"),
         os.path.join(HERE, "host", "irbc_main.cpp"), "-L" + HERE, "-lirbc_host", "-lirbc_amd",
         "-Wl,-rpath,$ORIGIN"])


def build_oracle():
    run(["make", "-C", os.path.join(REPO, "oracle"), "-j8"])


def build_all():
    build_lib()
    build_probe()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "host": build_host, "oracle": build_oracle,
         "probe": build_probe}[w]()
