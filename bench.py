#!/usr/bin/env python3
"""Benchmark of the MI355X Gaussian FIR path (dip::Gauss, method "FIR").

Headline workload: an 8192x8192 uint8 image (uniform 0..255, seeded) filtered to
sfloat with sigma in {1, 4, 16} (isotropic) and derivative orders (0,0), (1,0),
(0,2), boundary "mirror", truncation 3.  One step = those nine dip::Gauss calls
on the HBM-resident input, in the default exact mode (bit-identical to the
reference).  `value` = output pixels per second over the step (GPix/s).

The JSON line also carries: per-case rates with HBM and VALU roofline fractions,
the fast (FMA) mode, the 3-D anisotropic volume, the batched uneven form against
its per-image loop, forward+backward, the dominant kernel's roofline (HIP events
on the launch stream over the timed region) and the reference's own CPU path
timed on this host.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--no-cpu] [--quick]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9  # CUs x SIMDs x lanes x clock: FP32 lane-ops/s without packed math
SIGMAS = (1.0, 4.0, 16.0)
ORDERS = ((0, 0), (1, 0), (0, 2))
SIZE = 8192


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def radius(s, o, dbl=False):
    import diplib_amd as dip
    return dip.gauss_radius(s, o, 3.0, dbl)


def valu_ops_per_px(sig, ords, fast):
    per = 0
    for s, o in zip(sig, ords):
        r = radius(s, o)
        per += (2 * r + 1) if fast else (3 * r + 1)
    return per


def read_trace(L):
    from diplib_amd._lib import DipamdTraceRecord
    cap = 1 << 16
    recs = (DipamdTraceRecord * cap)()
    n = L.dipamd_trace_read(recs, cap)
    return [(recs[i].name.decode(), recs[i].ms, recs[i].bytes) for i in range(n)]


def cpu_baseline(x_np, budget_s=25.0):
    """The reference's own dip::Gauss (oracle/_ref/libdipref.so, built from the
    reference sources) with OpenMP on this host's cores; falls back to the C oracle
    (a restatement) when the reference library is not present."""
    import numpy as np
    import oracle_ctypes as oc
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    use_ref = oc.ref_available()
    kind = "reference" if use_ref else "port"
    done_px = 0
    t_total = 0.0
    ncalls = 0
    t_start = time.perf_counter()
    for s in SIGMAS:
        for o in ORDERS:
            t0 = time.perf_counter()
            oc.gauss_fir(x_np, [s], list(o), ["mirror"], 3.0, nthreads=cores, use_ref=use_ref)
            t_total += time.perf_counter() - t0
            done_px += x_np.size
            ncalls += 1
            if time.perf_counter() - t_start > budget_s:
                break
        if time.perf_counter() - t_start > budget_s:
            break
    gpix = done_px / t_total / 1e9
    return {"value": round(gpix, 4), "unit": "GPix/s", "cores": cores, "kind": kind,
            "sample": f"{ncalls} of the 9 dip::Gauss calls of the 8192x8192 uint8 workload, one pass each "
                      f"({t_total:.2f} s of CPU work, OpenMP {cores} threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; the multi-GPU (sharded) path is a design only "
              "-- run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cases
    import diplib_amd as dip
    from diplib_amd import _lib

    L = _lib.lib()
    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}; library {L.dipamd_version().decode()}")

    x_np = cases.make_input((SIZE, SIZE), "uint8", seed=2)
    x = torch.from_numpy(x_np).to(dev)
    outs = [torch.empty((SIZE, SIZE), dtype=torch.float32, device=dev) for _ in range(len(SIGMAS) * len(ORDERS))]
    combos = [(s, o) for s in SIGMAS for o in ORDERS]

    def step(fast=False):
        for i, (s, o) in enumerate(combos):
            dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    L.dipamd_trace_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    trace = read_trace(L)
    L.dipamd_trace_enable(0)
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    px_per_step = len(combos) * SIZE * SIZE
    value = px_per_step / (ms_per_step * 1e-3) / 1e9
    log(f"headline: {ms_per_step:.3f} ms/step, {value:.2f} GPix/s")

    # dominant kernel over the timed region
    agg = {}
    for name, ms, by in trace:
        e = agg.setdefault(name, [0.0, 0, 0])
        e[0] += ms
        e[1] += 1
        e[2] += by
    kernels = {}
    for name, (tot, cnt, by) in agg.items():
        avg_ms = tot / cnt
        gbs = (by / cnt) / (avg_ms * 1e-3) / 1e9
        kernels[name] = {"launches": cnt, "total_ms": round(tot, 4), "avg_ms": round(avg_ms, 5),
                         "bytes_per_launch": by // cnt, "achieved_GBs": round(gbs, 1),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(agg.items(), key=lambda kv: kv[1][0])[0] if agg else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if dom and os.path.exists(tfile):
        try:
            with open(tfile) as f:
                traffic = json.load(f).get(dom, {}).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = None
    if dom:
        k = kernels[dom]
        roofline = {"bound": "hbm", "achieved": k["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": k["hbm_frac"], "traffic": traffic, "kernel": dom,
                    "bytes_per_launch": k["bytes_per_launch"], "avg_launch_ms": k["avg_ms"]}

    # per-case rates (event-timed) and roofline fractions
    per_case = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for fast in (False, True):
        for i, (s, o) in enumerate(combos):
            for _ in range(2):
                dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)
            reps = 5
            ev0.record()
            for _ in range(reps):
                dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / reps
            gbs = SIZE * SIZE * 5 / (ms * 1e-3) / 1e9
            ops = SIZE * SIZE * valu_ops_per_px([s, s], list(o), fast)
            per_case.append({"sigma": s, "orders": list(o), "mode": "fast" if fast else "exact",
                             "ms": round(ms, 4), "GPix_s": round(SIZE * SIZE / (ms * 1e-3) / 1e9, 2),
                             "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                             "valu_frac": round(ops / (ms * 1e-3) / VALU_PEAK_OPS, 4)})
    fast_ms = sum(c["ms"] for c in per_case if c["mode"] == "fast")
    fast_value = px_per_step / (fast_ms * 1e-3) / 1e9

    # exactness of the timed outputs: the sigma=1, order (1,0) output against the reference's hash
    import hashlib
    with open(os.path.join(ROOT, "tests", "golden", "large_cases.json")) as f:
        gold = json.load(f)
    outs[1].zero_()
    dip.GaussFIR(x, [1.0], [1, 0], ["mirror"], 3.0, out=outs[1])
    torch.cuda.synchronize()
    bitexact = hashlib.sha256(outs[1].cpu().numpy().tobytes()).hexdigest() == gold["u8_8192sq_s1_o10"]["output_sha256"]

    extra = {}
    if not a.quick:
        extra.update(bench_volume(dip, torch, cases, dev))
        extra.update(bench_batched(dip, torch, np, dev, L))
        extra.update(bench_backward(dip, torch, dev))
    del outs
    cpu = None
    if not a.no_cpu:
        log("CPU baseline (reference dip::Gauss, OpenMP)")
        cpu = cpu_baseline(x_np)
        log(f"CPU baseline: {cpu['value']} GPix/s on {cpu['cores']} cores ({cpu['kind']})")

    line = {
        "metric": "GPix/s + % HBM roofline: Gauss fwd, batched, fwd+bwd vs plain GPU and CPU",
        "value": round(value, 3), "unit": "GPix/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "dip::Gauss FIR, 8192x8192 uint8 -> sfloat, sigma {1,4,16} x orders "
                               "{(0,0),(1,0),(0,2)}, mirror, truncation 3 (9 calls per step)",
                   "image": [SIZE, SIZE], "in_dtype": "uint8", "out_dtype": "sfloat", "mode": "exact"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bitexact_vs_reference": bitexact,
        "fast_mode_GPix_s": round(fast_value, 3),
        "cases": per_case,
        "kernels": kernels,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


def bench_volume(dip, torch, cases, dev):
    out = {}
    try:
        xv = torch.from_numpy(cases.make_input((256, 512, 512), "float32", 3, 0.01)).to(dev)
        yv = torch.empty_like(xv)
        res = {}
        for bc in ("mirror", "periodic"):
            for fast in (False, True):
                for _ in range(2):
                    dip.GaussFIR(xv, [2.0, 2.0, 1.0], [0], [bc], 3.0, out=yv, fast=fast)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(5):
                    dip.GaussFIR(xv, [2.0, 2.0, 1.0], [0], [bc], 3.0, out=yv, fast=fast)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 5
                n = xv.numel()
                res[f"{bc}_{'fast' if fast else 'exact'}"] = {
                    "ms": round(ms, 4), "GPix_s": round(n / (ms * 1e-3) / 1e9, 2),
                    "hbm_frac": round(n * 8 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        out["volume_512x512x256_s221"] = res
        log(f"3-D volume: {res}")
    except Exception as e:  # report, do not hide
        out["volume_512x512x256_s221"] = {"error": repr(e)}
    return out


def bench_batched(dip, torch, np, dev, L):
    out = {}
    if not hasattr(dip, "GaussBatch"):
        return out
    try:
        rng = np.random.default_rng(1234)
        n = 512
        ws = rng.integers(64, 2049, n)
        hs = rng.integers(64, 2049, n)
        sig = np.exp(rng.uniform(np.log(0.5), np.log(8.0), n))
        g = torch.Generator(device=dev)
        g.manual_seed(99)
        ins = [torch.rand((int(h), int(w)), generator=g, device=dev, dtype=torch.float32) for w, h in zip(ws, hs)]
        outs_b = [torch.empty_like(t) for t in ins]
        outs_l = [torch.empty_like(t) for t in ins]
        px = int(sum(int(w) * int(h) for w, h in zip(ws, hs)))
        plan = dip.GaussBatch(ins, outs_b, [(float(s), float(s)) for s in sig], boundary="mirror")
        loop = dip.GaussBatch(ins, outs_l, [(float(s), float(s)) for s in sig], boundary="mirror", tile_config=1)
        for _ in range(2):
            plan.run_graph()
            loop.run_per_image()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            plan.run_graph()
        e1.record()
        torch.cuda.synchronize()
        ms_b = e0.elapsed_time(e1) / reps
        e0.record()
        for _ in range(reps):
            loop.run_per_image()
        e1.record()
        torch.cuda.synchronize()
        ms_l = e0.elapsed_time(e1) / reps
        same = all(torch.equal(a_, b_) for a_, b_ in zip(outs_b, outs_l))
        out["batched_uneven_512"] = {
            "images": n, "pixels": px, "tile_config": plan.config, "work_items": plan.ntiles,
            "batched_ms": round(ms_b, 4), "batched_GPix_s": round(px / (ms_b * 1e-3) / 1e9, 2),
            "batched_hbm_frac": round(px * 8 / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "per_image_loop_ms": round(ms_l, 4), "per_image_loop_GPix_s": round(px / (ms_l * 1e-3) / 1e9, 2),
            "speedup_vs_loop": round(ms_l / ms_b, 3), "bitexact_vs_loop": bool(same)}
        log(f"batched: {out['batched_uneven_512']}")
        plan.close()
        loop.close()
    except Exception as e:
        out["batched_uneven_512"] = {"error": repr(e)}
    return out


def bench_backward(dip, torch, dev):
    out = {}
    if not hasattr(dip, "gauss_fir_autograd"):
        return out
    try:
        g = torch.Generator(device=dev)
        g.manual_seed(5)
        x = torch.rand((32, 1024, 1024), generator=g, device=dev, dtype=torch.float32)
        w = torch.randn((32, 1024, 1024), generator=g, device=dev, dtype=torch.float32)
        res = {}
        for order in (0, 1):
            sigma = torch.tensor(3.0, device=dev, dtype=torch.float64, requires_grad=True)
            xr = x.clone().requires_grad_(True)

            def fwd():
                return dip.gauss_fir_autograd(xr, sigma, [order, 0, 0], ["mirror"], 3.0, sigma_dims=(0, 1))

            def fwdbwd():
                y = fwd()
                y.backward(w)

            for _ in range(2):
                fwd()
                fwdbwd()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                with torch.no_grad():
                    fwd()
            e1.record()
            torch.cuda.synchronize()
            ms_f = e0.elapsed_time(e1) / 5
            e0.record()
            for _ in range(5):
                fwdbwd()
            e1.record()
            torch.cuda.synchronize()
            ms_fb = e0.elapsed_time(e1) / 5
            res[f"order{order}"] = {"fwd_ms": round(ms_f, 4), "fwd_bwd_ms": round(ms_fb, 4),
                                    "ratio": round(ms_fb / ms_f, 3),
                                    "fwd_GPix_s": round(x.numel() / (ms_f * 1e-3) / 1e9, 2)}
        out["fwd_bwd_32x1024x1024_s3"] = res
        log(f"fwd+bwd: {res}")
    except Exception as e:
        out["fwd_bwd_32x1024x1024_s3"] = {"error": repr(e)}
    return out


if __name__ == "__main__":
    main()
