"""Builds the native parts in-tree (no JIT cache, so the built files travel
with the source tree):

* ``kites_amd/libkites_amd.so``  -- the C ABI (capi.hip, batch_capi.hip,
  mlp.hip, ingest_capi.hip, ordered_capi.hip, noalloc_capi.hip + the
  header-only engine,
  compiled for gfx950),
  the host graph generators (gen.cpp) and graph files (graph_files.cpp).
* ``examples/_build/*``          -- the reference example rewritten against
  the drop-in C++ API, compiled for gfx950.

Run ``python -m kites_amd.build`` (``__graft_entry__.build()`` calls it).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkites_amd.so")

ARCH = os.environ.get("KITES_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay disabled in every HIP object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _jobs():
    try:
        return max(1, min(8, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def _parallel(cmds):
    """Runs independent compile commands concurrently (each its own process)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    cmds, host_objs, objs = [], [], []
    for src in ("gen.cpp", "graph_files.cpp"):
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmds.append(["g++", "-O3", "-fopenmp", "-fPIC", "-std=c++17", "-c",
                     os.path.join(CSRC, src), "-o", o])
        host_objs.append(o)
    for src in ("capi.hip", "batch_capi.hip", "mlp.hip", "ingest_capi.hip", "ordered_capi.hip",
                "noalloc_capi.hip"):
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([hipcc()] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    _parallel(cmds)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + host_objs +
         ["-lgomp", "-o", tmp])
    os.replace(tmp, LIB)


def build_examples():
    src_dir = os.path.join(ROOT, "examples")
    out_dir = os.path.join(src_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    exes = [os.path.join(out_dir, f[:-4]) for f in sorted(os.listdir(src_dir)) if f.endswith(".hip")]
    _parallel([[hipcc()] + HIP_FLAGS + ["-I", CSRC, e.replace("_build" + os.sep, "") + ".hip",
                                        "-o", e + ".tmp"] for e in exes])
    for e in exes:
        os.replace(e + ".tmp", e)


def build():
    build_lib()
    build_examples()


if __name__ == "__main__":
    build()
    sys.exit(0)
