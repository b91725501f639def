"""Build the MI355X engine libraries in-tree with hipcc.

    python -m moonlab_amd.build            # compile for gfx950

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instruction selection disabled.  Objects go to moonlab_amd/_build/ and the
shared libraries to moonlab_amd/lib/:

  libmoonlab_mi355x.so         the engine plus its own job / circuit API
                               (a drop-in for the reference library's
                               scheduler and QGTL entry points);
  libmoonlab_mi355x_plugin.so  the engine alone, for a process that already
                               runs the reference's scheduler: its job-API
                               references (moonlab_job_num_*, moonlab_job_to_json,
                               moonlab_register_backend) resolve to the host
                               library, so moonlab_mi355x_register_backend()
                               registers "mi355x" in the host's own registry.

Both link with -Bsymbolic, so calls inside a library bind to its own
definitions whatever else the process has loaded.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libmoonlab_mi355x.so")
PLUGIN = os.path.join(LIBDIR, "libmoonlab_mi355x_plugin.so")
ROUTE = os.path.join(LIBDIR, "libmoonlab_mi355x_route.so")
# translation units that implement the job / circuit API (left out of the plugin)
HOST_API = ("scheduler.cpp", "qgtl_circuit.cpp", "control_frame.cpp")
# the reference's GPU seam (quantum_state_create_gpu, qsim_gpu_route_*): its own
# library over the plugin, so only a process that links it gets those symbols
ROUTE_SRC = "gpu_route.cpp"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


# Per-file device flags.  sv_regpass.hip keeps 16 amplitudes per thread in a
# register array whose op arms differ only in register indices; sinking the
# arms' common tails merges them behind address phis that keep the array in
# scratch memory, so that pass is off there.
FILE_FLAGS = {"sv_regpass.hip": ["-mllvm", "-simplifycfg-sink-common=false"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + FILE_FLAGS.get(os.path.basename(src), []) + ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    route_obj = [o for o in objs if os.path.basename(o)[:-2] == ROUTE_SRC]
    objs = [o for o in objs if os.path.basename(o)[:-2] != ROUTE_SRC]
    plugin_objs = [o for o in objs if not os.path.basename(o)[:-2] in HOST_API]
    links = ((LIB, objs, []), (PLUGIN, plugin_objs, []),
             (ROUTE, route_obj, ["-L" + LIBDIR, "-lmoonlab_mi355x_plugin", "-Wl,-rpath,$ORIGIN"]))
    for out, parts, extra in links:
        tmp = out + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp] + parts + extra + \
              ["-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link of {os.path.basename(out)} failed:\n{r.stderr}")
        os.replace(tmp, out)
    if verbose:
        print(f"built {LIB} from {len(objs)} sources, {os.path.basename(PLUGIN)} from {len(plugin_objs)} and "
              f"{os.path.basename(ROUTE)}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
