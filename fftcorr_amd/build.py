"""Builds libfftcorr_hip.so for gfx950 from fftcorr_amd/csrc with hipcc.

Each .hip source compiles to an object in parallel, then one shared library
is linked in-tree at fftcorr_amd/_lib/libfftcorr_hip.so.

    python -m fftcorr_amd.build [--jobs N] [--incremental]

--incremental recompiles only sources older than their object or than any
header of csrc/ or include/ (a header change rebuilds everything).
"""

import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_lib", "obj")
LIB = os.path.join(HERE, "_lib", "libfftcorr_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FFTCORR_ARCH", "gfx950")

# -packed-fp32-ops off: hipcc then emits no v_pk_{add,mul,fma}_f32.
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _obj(src):
    return os.path.join(OBJDIR, os.path.basename(src) + ".o")


def _stale(src):
    obj = _obj(src)
    if not os.path.exists(obj):
        return True
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(ROOT, "include")
    headers += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in headers])
    return os.path.getmtime(obj) < newest


def _compile(src):
    obj = _obj(src)
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(jobs=None, verbose=True, incremental=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if _stale(s)] if incremental else srcs
    jobs = jobs or min(max(len(todo), 1), os.cpu_count() or 4, 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [_obj(s) for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    try:
        build(a.jobs, incremental=a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
