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
on the launch stream around every kernel of a repeat of the timed steps; the
headline region itself runs without them) and the reference's own CPU path
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
    return [(recs[i].name.decode(), recs[i].ms, recs[i].bytes, recs[i].ops) for i in range(n)]


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
    ap.add_argument("--case", nargs=2, metavar=("SIGMA", "ORDERS"),
                    help="profiling mode: run one case of the sweep (e.g. 16 00; or 'vol mirror' for the "
                         "512x512x256 sfloat volume, sigma (2,2,1); or 'iir 00' for dip::GaussIIR sigma 16, "
                         "'ft 00' for dip::GaussFT sigma 0.5, on the 8192^2 input; or 'bwd mirror' for "
                         "forward + backward through the autograd function, 32 sfloat 1024^2 planes, sigma 3 "
                         "on x and y, order 0; or 'deriv laplace' etc. for the derivative family at sigma 1 "
                         "and 4) --steps times, then a 256 MiB device copy as the calibration "
                         "kernel for the traffic counters")
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

    if a.case and a.case[0] == "bwd":
        g = torch.Generator(device=dev)
        g.manual_seed(5)
        xb = torch.rand((32, 1024, 1024), generator=g, device=dev, dtype=torch.float32).requires_grad_(True)
        wb = torch.randn((32, 1024, 1024), generator=g, device=dev, dtype=torch.float32)
        sb = torch.tensor(3.0, device=dev, dtype=torch.float64, requires_grad=True)
        for _ in range(a.steps):
            xb.grad = None
            sb.grad = None
            dip.gauss_fir_autograd(xb, sb, [0, 0, 0], [a.case[1]], 3.0, sigma_dims=(0, 1), sigma_host=3.0).backward(wb)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(a.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"case backward {a.case[1]}: {a.steps} forward + backward calls, then {a.steps} copies of 256 MiB")
        return
    if a.case and a.case[0] == "deriv":
        # the derivative family on the 8192^2 uint8 input, sigma 1 then 4 (FIR)
        fn = {"gradient": dip.Gradient, "gradmag": dip.GradientMagnitude, "hessian": dip.Hessian,
              "laplace": dip.Laplace, "dgg": dip.Dgg}[a.case[1]]
        xd = torch.from_numpy(cases.make_input((SIZE, SIZE), "uint8", seed=2)).to(dev)
        for s in (1.0, 4.0):
            for _ in range(a.steps):
                y = fn(xd, [s], "FIR", ["mirror"])
            torch.cuda.synchronize()
            del y
        log(f"case derivative {a.case[1]}: {a.steps} calls at sigma 1, then {a.steps} at sigma 4")
        return
    if a.case and a.case[0] == "vol":
        xv = torch.from_numpy(cases.make_input((256, 512, 512), "float32", 3, 0.01)).to(dev)
        yv = torch.empty_like(xv)
        for _ in range(a.steps):
            dip.GaussFIR(xv, [2.0, 2.0, 1.0], [0], [a.case[1]], 3.0, out=yv)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(a.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"case volume {a.case[1]}: {a.steps} calls, then {a.steps} copies of 256 MiB")
        return
    x_np = cases.make_input((SIZE, SIZE), "uint8", seed=2)
    x = torch.from_numpy(x_np).to(dev)
    if a.case:
        kind = a.case[0] if a.case[0] in ("iir", "ft") else "fir"
        s = {"iir": 16.0, "ft": 0.5}.get(kind) or float(a.case[0])
        o = [int(c) for c in a.case[1]]
        y = torch.empty((SIZE, SIZE), dtype=torch.float32, device=dev)
        for _ in range(a.steps):
            if kind == "iir":
                dip.GaussIIR(x, [s], o, ["mirror"], out=y)
            elif kind == "ft":
                dip.GaussFT(x, [s], o, 3.0, "spatial", "spatial", ["mirror"], out=y)
            else:
                dip.GaussFIR(x, [s], o, ["mirror"], 3.0, out=y)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB read + 256 MiB written
        dst = torch.empty_like(src)
        for _ in range(a.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"case {kind} sigma={s} orders={o}: {a.steps} calls, then {a.steps} copies of 256 MiB")
        return
    outs = [torch.empty((SIZE, SIZE), dtype=torch.float32, device=dev) for _ in range(len(SIGMAS) * len(ORDERS))]
    combos = [(s, o) for s in SIGMAS for o in ORDERS]

    def step(fast=False):
        for i, (s, o) in enumerate(combos):
            dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the same K steps again with HIP events recorded on the launch stream around every
    # kernel: per-kernel launch durations for the roofline (the events themselves add
    # about 4 % to a step, so the headline region above runs without them)
    L.dipamd_trace_enable(1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    trace = read_trace(L)
    L.dipamd_trace_enable(0)
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    ms_per_step_traced = (t3 - t2) * 1e3 / a.steps
    px_per_step = len(combos) * SIZE * SIZE
    value = px_per_step / (ms_per_step * 1e-3) / 1e9
    log(f"headline: {ms_per_step:.3f} ms/step, {value:.2f} GPix/s")

    copy_roof = measure_copy_roof(L, torch, dev)
    log(f"copy roof: {copy_roof}")

    # dominant kernel over the timed region
    agg = {}
    for name, ms, by, ops in trace:
        e = agg.setdefault(name, [0.0, 0, 0, 0])
        e[0] += ms
        e[1] += 1
        e[2] += by
        e[3] += ops
    kernels = {}
    for name, (tot, cnt, by, ops) in agg.items():
        avg_ms = tot / cnt
        gbs = (by / cnt) / (avg_ms * 1e-3) / 1e9
        kernels[name] = {"launches": cnt, "total_ms": round(tot, 4), "avg_ms": round(avg_ms, 5),
                         "bytes_per_launch": by // cnt, "achieved_GBs": round(gbs, 1),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                         "copy_roof_frac": round(gbs / roof_for(copy_roof, by / cnt / (SIZE * SIZE)), 4),
                         "ops_per_launch": ops // cnt,
                         "valu_frac": round((ops / cnt) / (avg_ms * 1e-3) / VALU_PEAK_OPS, 4)}
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
        hbm = {"achieved": k["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k["hbm_frac"],
               "bytes_per_launch": k["bytes_per_launch"],
               "measured_copy_GBs": roof_for(copy_roof, k["bytes_per_launch"] / (SIZE * SIZE)),
               "frac_of_measured_copy": k["copy_roof_frac"]}
        # exact-mode arithmetic is one FP32 add or multiply per lane-op (no FMA, no packed
        # math), so its ceiling is the unpacked VALU rate; the published FP32 peak counts
        # an FMA as two FLOPs and packed pairs twice (4x this rate)
        valu = {"achieved": round(k["ops_per_launch"] / (k["avg_ms"] * 1e-3) / 1e12, 3),
                "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "TFLOP/s", "frac": k["valu_frac"],
                "ops_per_launch": k["ops_per_launch"], "published_peak": 157.3,
                "frac_of_published": round(k["ops_per_launch"] / (k["avg_ms"] * 1e-3) / 157.3e12, 4)}
        bound = "valu" if k["valu_frac"] > k["hbm_frac"] else "hbm"
        top = valu if bound == "valu" else hbm
        roofline = {"bound": bound, "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"],
                    "frac": top["frac"], "traffic": traffic, "kernel": dom,
                    "bytes_per_launch": k["bytes_per_launch"], "avg_launch_ms": k["avg_ms"],
                    "hbm": hbm, "valu": valu}

    # per-case rates (event-timed) and roofline fractions; each case also once more with
    # the kernel trace on, whose per-call kernel time (kernel_ms) separates the kernels
    # from the launch gaps between calls
    per_case = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for fast in (False, True):
        for i, (s, o) in enumerate(combos):
            for _ in range(3):
                dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)
            reps = 20
            torch.cuda.synchronize()
            ev0.record()
            for _ in range(reps):
                dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / reps
            L.dipamd_trace_enable(1)
            for _ in range(reps):
                dip.GaussFIR(x, [s], list(o), ["mirror"], 3.0, out=outs[i], fast=fast)
            torch.cuda.synchronize()
            kms = sum(t[1] for t in read_trace(L)) / reps
            L.dipamd_trace_enable(0)
            gbs = SIZE * SIZE * 5 / (ms * 1e-3) / 1e9
            ops = SIZE * SIZE * valu_ops_per_px([s, s], list(o), fast)
            per_case.append({"sigma": s, "orders": list(o), "mode": "fast" if fast else "exact",
                             "ms": round(ms, 4), "kernel_ms": round(kms, 4),
                             "GPix_s": round(SIZE * SIZE / (ms * 1e-3) / 1e9, 2),
                             "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                             "copy_roof_frac": round(gbs / copy_roof["u8_to_f32_GBs"], 4),
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
        extra.update(bench_derivatives(dip, torch, x))
        extra.update(bench_iir(dip, torch, np, x, x_np, not a.no_cpu))
        extra.update(bench_fp32_vs_fp64(dip, torch, np, cases, dev))
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
        "ms_per_step_traced": round(ms_per_step_traced, 4),
        "copy_roof": copy_roof,
        "cpu_baseline": cpu,
        "bitexact_vs_reference": bitexact,
        "fast_mode_GPix_s": round(fast_value, 3),
        "cases": per_case,
        "kernels": kernels,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


def measure_copy_roof(L, torch, dev, reps=20):
    """What the memory system sustains for the FIR path's two traffic shapes, measured
    with the library's streaming copy kernels (k_roof.hip, interleaved layout: every
    load and store instruction of a wave covers one contiguous span) on 8192^2 samples:
    uint8 -> sfloat (5 bytes per sample) and sfloat -> sfloat (8)."""
    n = SIZE * SIZE
    out = {"samples": n, "kernel": "k_copy_u8_f32_il / k_copy_f32_f32_il"}
    st = torch.cuda.current_stream(dev).cuda_stream
    dst = torch.empty(n, dtype=torch.float32, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    from diplib_amd import _lib
    for key, src, dt in (("u8_to_f32", torch.randint(0, 256, (n,), dtype=torch.uint8, device=dev), 0),
                         ("f32_to_f32", torch.rand(n, device=dev), 1)):
        code = _lib.DT_UINT8 if dt == 0 else _lib.DT_SFLOAT
        for _ in range(3):
            if L.dipamd_measure_copy(src.data_ptr(), code, dst.data_ptr(), n, 2, st) != 0:
                raise RuntimeError("copy kernel failed")
        e0.record()
        for _ in range(reps):
            L.dipamd_measure_copy(src.data_ptr(), code, dst.data_ptr(), n, 2, st)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[key + "_us"] = round(ms * 1e3, 2)
        out[key + "_GBs"] = round(n * (src.element_size() + 4) / (ms * 1e-3) / 1e9, 1)
        ok = torch.equal(dst, src.float()) if dt == 0 else torch.equal(dst, src)
        if not ok:
            raise RuntimeError("copy kernel result wrong")
    return out


def roof_for(copy_roof, bytes_per_px):
    """The measured copy roof of the traffic shape nearest to bytes_per_px (5: uint8 in,
    sfloat out; 8: sfloat in and out)."""
    return copy_roof["u8_to_f32_GBs"] if bytes_per_px < 6.5 else copy_roof["f32_to_f32_GBs"]


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


def err_vs_fp64(got, ref64):
    """FP32 result against an FP64 result: max|a-b| / max|b| and DIPlib's
    MaximumRelativeError, max |a-b| / (|a|+|b|)/2 with 0 where both are 0
    (error.cpp:77-109; ill-conditioned at zero crossings of derivative outputs)."""
    import numpy as np
    a = got.astype(np.float64)
    d = np.abs(a - ref64)
    den = 0.5 * (np.abs(a) + np.abs(ref64))
    mre = float(np.max(np.where(den > 0, d / np.where(den > 0, den, 1.0), 0.0))) if d.size else 0.0
    scale = float(np.max(np.abs(ref64))) if d.size else 0.0
    return {"max_abs_over_max_ref": float(d.max() / scale) if scale > 0 else 0.0, "dip_max_rel_err": mre}


def bench_fp32_vs_fp64(dip, torch, np, cases, dev):
    """FP32 outputs (exact mode, so the reference's own sfloat results) against the
    FP64 oracle, for every workload.  Large images are checked on bands of rows (or
    planes) extended by the filter radius, whose centre rows do not see the band's
    cut: the top band also covers the image edge's boundary extension."""
    import oracle_ctypes as oc
    res = {}
    nthr = min(16, len(os.sched_getaffinity(0)))

    def ref64(x, *args):
        return oc.gauss_fir(x, *args, nthreads=nthr)

    try:
        band = 64
        # 512x512 sfloat, sigma 2 (the reference's CPU-runnable case): whole image
        x = cases.make_input((512, 512), "float32", 1)
        y = dip.GaussFIR(torch.from_numpy(x).to(dev), [2.0], [0], ["mirror"], 3.0).cpu().numpy()
        res["f32_512x512_s2"] = err_vs_fp64(y, ref64(x.astype(np.float64), [2.0], [0], ["mirror"], 3.0))
        # 8192^2 uint8 sweep: a band at the top edge and one at the centre
        xu = cases.make_input((SIZE, SIZE), "uint8", 2)
        xd = torch.from_numpy(xu).to(dev)
        for s in SIGMAS:
            for o in ORDERS:
                y = dip.GaussFIR(xd, [s], list(o), ["mirror"], 3.0)
                r = radius(s, o[1])
                worst = {"max_abs_over_max_ref": 0.0, "dip_max_rel_err": 0.0}
                for y0 in (0, SIZE // 2):
                    lo = max(0, y0 - r)
                    ref = ref64(xu[lo:y0 + band + r].astype(np.float64), [s], list(o), ["mirror"], 3.0)
                    e = err_vs_fp64(y[y0:y0 + band].cpu().numpy(), ref[y0 - lo:y0 - lo + band])
                    worst = {k: max(worst[k], e[k]) for k in worst}
                res[f"u8_8192sq_s{int(s)}_o{o[0]}{o[1]}"] = worst
        del xd
        # 3-D volume: a slab of planes at the centre and at the first plane
        xv = cases.make_input((256, 512, 512), "float32", 3, 0.01)
        xvd = torch.from_numpy(xv).to(dev)
        rz = radius(1.0, 0)
        for bc in ("mirror", "periodic"):
            yv = dip.GaussFIR(xvd, [2.0, 2.0, 1.0], [0], [bc], 3.0)
            worst = {"max_abs_over_max_ref": 0.0, "dip_max_rel_err": 0.0}
            for z0 in ((0, 128) if bc == "mirror" else (128,)):
                lo = max(0, z0 - rz)
                ref = ref64(xv[lo:z0 + 8 + rz].astype(np.float64), [2.0, 2.0, 1.0], [0], [bc], 3.0)
                e = err_vs_fp64(yv[z0:z0 + 8].cpu().numpy(), ref[z0 - lo:z0 - lo + 8])
                worst = {k: max(worst[k], e[k]) for k in worst}
            res[f"f32_512x512x256_s221_{bc}"] = worst
        del xvd
        # FP64 volume: the device's dfloat path against the FP64 oracle (whole volume)
        x64 = cases.make_input((128, 256, 256), "float64", 4, 0.01)
        y64 = dip.GaussFIR(torch.from_numpy(x64).to(dev), [2.0, 2.0, 1.0], [0], ["mirror"], 3.0).cpu().numpy()
        res["f64_256x256x128_s221_mirror"] = err_vs_fp64(y64, ref64(x64, [2.0, 2.0, 1.0], [0], ["mirror"], 3.0))
        # batched uneven form: the first 8 images of the benchmark's batch
        rng = np.random.default_rng(1234)
        ws, hs = rng.integers(64, 2049, 512), rng.integers(64, 2049, 512)
        sig = np.exp(rng.uniform(np.log(0.5), np.log(8.0), 512))
        ins = [torch.from_numpy(cases.make_input((int(h), int(w)), "float32", 50 + i)).to(dev)
               for i, (w, h) in enumerate(zip(ws[:8], hs[:8]))]
        outs = [torch.empty_like(t) for t in ins]
        plan = dip.GaussBatch(ins, outs, [(float(s), float(s)) for s in sig[:8]], boundary="mirror")
        plan.run()
        torch.cuda.synchronize()
        worst = {"max_abs_over_max_ref": 0.0, "dip_max_rel_err": 0.0}
        for t, o, s in zip(ins, outs, sig[:8]):
            e = err_vs_fp64(o.cpu().numpy(), ref64(t.cpu().numpy().astype(np.float64), [float(s)], [0],
                                                          ["mirror"], 3.0))
            worst = {k: max(worst[k], e[k]) for k in worst}
        plan.close()
        res["batched_uneven_first8"] = worst
        # differentiation form, forward: one 1024^2 image, sigma 3, orders 0 and 1
        x5 = cases.make_input((1024, 1024), "float32", 5)
        for order in (0, 1):
            y5 = dip.GaussFIR(torch.from_numpy(x5).to(dev), [3.0], [order, 0], ["mirror"], 3.0).cpu().numpy()
            res[f"f32_1024sq_s3_o{order}0"] = err_vs_fp64(
                y5, ref64(x5.astype(np.float64), [3.0], [order, 0], ["mirror"], 3.0))
        log(f"FP32 vs FP64: {res}")
    except Exception as e:  # report, do not hide
        res["error"] = repr(e)
    return {"fp32_vs_fp64": res}


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


def bench_derivatives(dip, torch, x):
    """The derivative family on the 8192^2 uint8 input (FIR): time per call and input
    pixels per second; each call is n Gaussian derivative passes (plus one element-wise
    pass for GradientMagnitude and Laplace)."""
    out = {}
    try:
        res = {}
        n = x.numel()
        for s in (1.0, 4.0):
            for name, fn in (("Gradient", dip.Gradient), ("Hessian", dip.Hessian), ("Laplace", dip.Laplace),
                             ("GradientMagnitude", dip.GradientMagnitude)):
                for _ in range(2):
                    y = fn(x, [s], "FIR", ["mirror"])
                torch.cuda.synchronize()
                del y
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(3):
                    y = fn(x, [s], "FIR", ["mirror"])
                e1.record()
                torch.cuda.synchronize()
                del y
                ms = e0.elapsed_time(e1) / 3
                res[f"{name}_s{int(s)}"] = {"ms": round(ms, 4), "GPix_s": round(n / (ms * 1e-3) / 1e9, 2)}
        out["derivative_family_8192sq_u8"] = res
        log(f"derivative family: {res}")
    except Exception as e:  # report, do not hide
        out["derivative_family_8192sq_u8"] = {"error": repr(e)}
    return out


def bench_iir(dip, torch, np, x, x_np, with_cpu):
    """dip::GaussIIR (what "best" picks for sigma > 10) on the 8192^2 uint8 input: time
    per call on the GPU; with the CPU leg, the reference's own GaussIIR on the same
    image with OpenMP on this host's cores, and the two outputs compared bit for bit."""
    out = {}
    try:
        res = {}
        n = x.numel()
        y = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        for s in (16.0, 64.0):
            for o in ((0, 0), (1, 0), (0, 2)):
                dip.GaussIIR(x, [s], list(o), ["mirror"], out=y)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(3):
                    dip.GaussIIR(x, [s], list(o), ["mirror"], out=y)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 3
                res[f"s{int(s)}_o{o[0]}{o[1]}"] = {"ms": round(ms, 4), "GPix_s": round(n / (ms * 1e-3) / 1e9, 2)}
        out["gauss_iir_8192sq_u8"] = res
        log(f"GaussIIR: {res}")
        if with_cpu:
            import oracle_ctypes as oc
            if oc.ref_available():
                try:
                    cores = len(os.sched_getaffinity(0))
                except AttributeError:
                    cores = os.cpu_count() or 1
                dip.GaussIIR(x, [16.0], [0, 0], ["mirror"], out=y)
                gpu = y.cpu().numpy()
                t0 = time.perf_counter()
                ref = oc.gauss_iir(x_np, [16.0], [0, 0], ["mirror"], nthreads=cores, use_ref=True)
                t = time.perf_counter() - t0
                out["gauss_iir_cpu_reference"] = {
                    "value": round(n / t / 1e9, 4), "unit": "GPix/s", "cores": cores, "kind": "reference",
                    "sample": "one dip::GaussIIR call, sigma 16, order (0,0), on the full 8192x8192 uint8 image",
                    "bitexact_vs_gpu": bool(np.array_equal(gpu, ref))}
                log(f"GaussIIR CPU reference: {out['gauss_iir_cpu_reference']}")
    except Exception as e:  # report, do not hide
        out["gauss_iir_8192sq_u8"] = {"error": repr(e)}
    return out


def bench_backward(dip, torch, dev):
    """Forward and forward+backward through the torch.autograd.Function (sigma
    requires grad; loss = sum(W * out) via out.backward(W)): a stack of 32 sfloat
    1024x1024 images (sigma shared by x and y) and a 128^3 sfloat volume (sigma shared
    by all three axes), derivative orders 0 and 1 along x.  x.grad is reset before
    each backward so no gradient accumulation is timed; sigma's host value is passed
    (sigma_host) so the device sigma is not read back (a stream wait) every call.
    The wall-clock ratio includes the Python and autograd overhead of both loops (the
    forward-only loop is partly bound by it); kernel_ratio is the same ratio over the
    library's kernel time alone (HIP events around every launch, in separate loops)."""
    out = {}
    if not hasattr(dip, "gauss_fir_autograd"):
        return out
    from diplib_amd import _lib
    L = _lib.lib()

    def kernel_ms(fn, n=5):
        torch.cuda.synchronize()
        L.dipamd_trace_enable(1)
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        tr = read_trace(L)
        L.dipamd_trace_enable(0)
        return sum(ms for _, ms, _, _ in tr) / n

    def run(shape, sdims, orders, seed):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        x = torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
        w = torch.randn(shape, generator=g, device=dev, dtype=torch.float32)
        sigma = torch.tensor(3.0, device=dev, dtype=torch.float64, requires_grad=True)
        xr = x.clone().requires_grad_(True)

        def fwd():
            return dip.gauss_fir_autograd(xr, sigma, orders, ["mirror"], 3.0, sigma_dims=sdims, sigma_host=3.0)

        def fwdbwd():
            xr.grad = None
            sigma.grad = None
            fwd().backward(w)

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

        def fwd_nograd():
            with torch.no_grad():
                fwd()

        k_f = kernel_ms(fwd_nograd)
        k_fb = kernel_ms(fwdbwd)
        return {"fwd_ms": round(ms_f, 4), "fwd_bwd_ms": round(ms_fb, 4), "ratio": round(ms_fb / ms_f, 3),
                "fwd_GPix_s": round(x.numel() / (ms_f * 1e-3) / 1e9, 2),
                "fwd_bwd_GPix_s": round(x.numel() / (ms_fb * 1e-3) / 1e9, 2),
                "fwd_kernel_ms": round(k_f, 4), "fwd_bwd_kernel_ms": round(k_fb, 4),
                "kernel_ratio": round(k_fb / k_f, 3)}

    for key, shape, sdims, nd in (("fwd_bwd_32x1024x1024_s3", (32, 1024, 1024), (0, 1), 3),
                                  ("fwd_bwd_128x128x128_s3", (128, 128, 128), (0, 1, 2), 3)):
        try:
            res = {}
            for order in (0, 1):
                res[f"order{order}"] = run(shape, sdims, [order] + [0] * (nd - 1), 5)
            out[key] = res
            log(f"{key}: {res}")
        except Exception as e:  # report, do not hide
            out[key] = {"error": repr(e)}
    return out


if __name__ == "__main__":
    main()
