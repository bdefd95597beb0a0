"""Build the native parts in-tree:

  cgbdsm_amd/libcgbdsm.so        HIP engine for gfx950 (hipcc)
  cgbdsm_amd/libcgbdsm_host.so   host pipeline (g++)
  cgbdsm_amd/bin/cgcsm_*         the reference-compatible CLI executables
  oracle/liboracle.so            CPU oracle (test infrastructure, plain C)

Objects go to cgbdsm_amd/_build/.  Every HIP file is compiled with packed
FP32 disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIP_DIR = os.path.join(PKG, "csrc", "hip")
HOST_DIR = os.path.join(PKG, "csrc", "host")
CLI_DIR = os.path.join(PKG, "csrc", "cli")
BUILD = os.path.join(PKG, "_build")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("CGBDSM_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", f"-I{INC}"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INC}", "-pthread"]
CLI_TARGETS = ["cgcsm_merged_matching", "cgcsm_gamma", "cgcsm_merged_matching_global", "cgcsm_gpma"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out)


def _srcs(d, ext):
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def build_host(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(PKG, "libcgbdsm_host.so")
    objs = []
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = []
        for s in _srcs(HOST_DIR, ".cpp"):
            o = os.path.join(BUILD, "host_" + os.path.basename(s) + ".o")
            objs.append(o)
            futs.append(ex.submit(_run, ["g++", *CXX_FLAGS, "-c", s, "-o", o], verbose))
        for f in futs:
            f.result()
    _run(["g++", "-shared", "-pthread", "-o", out, *objs], verbose)
    return objs


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(PKG, "libcgbdsm.so")
    objs = []
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = []
        for s in _srcs(HIP_DIR, ".hip"):
            o = os.path.join(BUILD, "hip_" + os.path.basename(s) + ".o")
            objs.append(o)
            futs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", s, "-o", o], verbose))
        for f in futs:
            f.result()
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out, *objs], verbose)
    return objs


def build_cli(host_objs, verbose=False):
    if not os.path.isdir(CLI_DIR):
        return
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    main = os.path.join(CLI_DIR, "main_cgcsm.cpp")
    obj = os.path.join(BUILD, "cli_main.o")
    _run(["g++", *CXX_FLAGS, f"-I{HOST_DIR}", "-c", main, "-o", obj], verbose)
    first = os.path.join(bindir, CLI_TARGETS[0])
    # The engine library is found next to the executables' parent directory.
    _run(["g++", "-pthread", "-o", first, obj, *host_objs, f"-L{PKG}", "-lcgbdsm",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,--disable-new-dtags"], verbose)
    for t in CLI_TARGETS[1:]:
        dst = os.path.join(bindir, t)
        if os.path.lexists(dst):
            os.remove(dst)
        os.symlink(CLI_TARGETS[0], dst)


def build_all(verbose=False):
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        f_or = ex.submit(build_oracle, verbose)
        f_host = ex.submit(build_host, verbose)
        f_hip = ex.submit(build_hip, verbose)
        f_or.result()
        host_objs = f_host.result()
        f_hip.result()
    build_cli(host_objs, verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", PKG)
