"""Builds diplib_amd/libdipamd.so (the C-ABI library: HIP kernels for gfx950 plus
the host planner) with hipcc, one translation unit per process, in parallel.

    python -m diplib_amd.build            # full rebuild
    python -m diplib_amd.build --changed  # only translation units older than their sources

Objects go to diplib_amd/_build/ (not committed); the library lands in the package
directory so it travels with the source tree.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdipamd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DIPAMD_ARCH", "gfx950")

# -ffp-contract=off: the exact mode must round like the reference (no FMA);
# packed FP32 ops are disabled for all device code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, f))
    return out


def headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h", ".inc"))] + \
        [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def _compile(src):
    obj = _obj(src)
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + FLAGS + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (p.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src, p.returncode, err


def build(changed_only=False, jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    newest_hdr = max(os.path.getmtime(h) for h in headers())
    todo = []
    for s in srcs:
        o = _obj(s)
        if changed_only and os.path.exists(o) and os.path.getmtime(o) > max(os.path.getmtime(s), newest_hdr):
            continue
        todo.append(s)
    jobs = jobs or min(len(todo) or 1, max(1, (os.cpu_count() or 4)))
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, err in ex.map(_compile, todo):
            if verbose:
                print(f"[build] {os.path.basename(src)}: {'ok' if rc == 0 else 'FAILED'}", flush=True)
            if rc != 0:
                failed.append((src, err))
            elif err and verbose:
                print(err, flush=True)
    if failed:
        for src, err in failed:
            print(f"--- {src}\n{err}", file=sys.stderr)
        raise RuntimeError("hipcc failed for: " + ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + [_obj(s) for s in srcs] + ["-ldl", "-lrocfft"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] linked {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--changed", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(changed_only=a.changed, jobs=a.j)
