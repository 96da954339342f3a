"""Deterministic synthetic inputs and the parity case lists shared by the golden
fixture generator (tests/golden/make_golden.py), the tests and the benchmark.

Images are numpy arrays in numpy axis order (..., y, x); sigma / order / boundary
lists are in dip axis order (x first), as dip::Gauss takes them.
"""
import numpy as np


def make_input(shape, dtype, seed, impulses=0.0):
    """Seeded input: uniform [0,1) for float types, uniform integers over the type's
    range (0..255 for uint8) for integer types, optional sparse impulses of amplitude 10
    (fraction `impulses` of the pixels, seeded positions)."""
    rng = np.random.default_rng(seed)
    dt = np.dtype(dtype)
    if dt.kind == "c":
        ft = np.float64 if dt == np.complex128 else np.float32
        return (rng.random(shape).astype(ft) + 1j * rng.random(shape).astype(ft)).astype(dt)
    if dt.kind == "f":
        x = rng.random(shape, dtype=np.float64).astype(dt) if dt == np.float64 else rng.random(shape, dtype=np.float32)
    elif dt == np.bool_:
        x = rng.random(shape) < 0.5
    else:
        info = np.iinfo(dt)
        lo, hi = (0, 256) if dt == np.uint8 else (max(info.min, -1000), min(info.max, 1000) + 1)
        x = rng.integers(lo, hi, size=shape, dtype=np.int64).astype(dt)
    if impulses > 0:
        n = x.size
        k = int(round(impulses * n))
        idx = rng.choice(n, size=k, replace=False)
        flat = x.reshape(-1)
        flat[idx] = flat[idx] + dt.type(10)
    return x


# Small full-output cases (pinned by fixtures written by the reference itself).
# (name, numpy shape, dtype, sigmas(dip order), orders, boundary list, out dtype or None)
SMALL_CASES = [
    ("f32_64x48_s2_o0_mirror", (48, 64), "float32", [2.0], [0], ["mirror"], None),
    ("f32_512x512_s2_o0_mirror", (512, 512), "float32", [2.0], [0], ["mirror"], None),
    ("u8_64x48_s1_o10_mirror", (48, 64), "uint8", [1.0], [1, 0], ["mirror"], None),
    ("u8_64x48_s1_o02_mirror", (48, 64), "uint8", [1.0], [0, 2], ["mirror"], None),
    ("u8_97x61_s4_o00_periodic", (61, 97), "uint8", [4.0], [0, 0], ["periodic"], None),
    ("f32_37x29x11_s221_mirror", (11, 29, 37), "float32", [2.0, 2.0, 1.0], [0], ["mirror"], None),
    ("f32_37x29x11_s221_periodic", (11, 29, 37), "float32", [2.0, 2.0, 1.0], [0], ["periodic"], None),
    ("f64_37x29x11_s221_mirror", (11, 29, 37), "float64", [2.0, 2.0, 1.0], [0], ["mirror"], None),
    ("f64_33x20_s3_o12_periodic", (20, 33), "float64", [3.0], [1, 2], ["periodic"], None),
    ("f32_130x70_s16_o00_mirror", (70, 130), "float32", [16.0], [0, 0], ["mirror"], None),
    ("f32_70x130_s8_o21_mirror", (130, 70), "float32", [8.0], [2, 1], ["mirror"], None),
    ("f32_41x23_s3_o3_mirror", (23, 41), "float32", [3.0], [3, 0], ["mirror"], None),
    ("f32_50x40_s05_mirror", (40, 50), "float32", [0.5], [0], ["mirror"], None),
    ("f32_50x40_aniso_mixedbc", (40, 50), "float32", [1.5, 2.5], [0, 1], ["periodic", "mirror"], None),
    ("f32_50x40_sigma0_x", (40, 50), "float32", [0.0, 2.0], [0], ["mirror"], None),
    ("f32_64x48_asym_mirror", (48, 64), "float32", [2.0], [1], ["asym mirror"], None),
    ("f32_64x48_asym_periodic", (48, 64), "float32", [2.0], [0, 1], ["asym periodic"], None),
    ("f32_64x48_add_zeros", (48, 64), "float32", [2.0], [0], ["add zeros"], None),
    ("f32_64x48_zero_order", (48, 64), "float32", [3.0], [0], ["zero order"], None),
    ("f32_64x48_add_max", (48, 64), "float32", [1.0], [0], ["add max"], None),
    ("u16_45x31_s2", (31, 45), "uint16", [2.0], [0], ["mirror"], None),
    ("s16_45x31_s2_o1", (31, 45), "int16", [2.0], [1], ["mirror"], None),
    ("s8_45x31_s2", (31, 45), "int8", [2.0], [0], ["periodic"], None),
    ("s32_45x31_s2", (31, 45), "int32", [2.0], [0], ["mirror"], None),
    ("u32_45x31_s2", (31, 45), "uint32", [2.0], [1], ["mirror"], None),
    ("bin_45x31_s2", (31, 45), "bool", [2.0], [0], ["mirror"], None),
    ("f32_1d_300_s5", (300,), "float32", [5.0], [0], ["mirror"], None),
    ("f32_4d_9x8x7x6", (6, 7, 8, 9), "float32", [1.0, 1.5, 1.0, 2.0], [0], ["mirror"], None),
    ("f32_out_u8_64x48", (48, 64), "float32", [2.0], [0], ["mirror"], "uint8"),
    ("u8_out_u16_64x48", (48, 64), "uint8", [2.0], [0], ["mirror"], "uint16"),
    ("f32_out_f64_64x48", (48, 64), "float32", [2.0], [1], ["mirror"], "float64"),
    ("f64_out_f32_64x48", (48, 64), "float64", [2.0], [0], ["periodic"], "float32"),
    ("f32_out_s16_64x48", (48, 64), "float32", [1.0], [1], ["mirror"], "int16"),
]
# Bounce cases: lines shorter than the kernel radius (mirror bounces, periodic wraps
# several times).  Widths 2..7 along x, heights 2..5 along y.
for w in range(2, 8):
    for h in (2, 3, 5):
        for bc in ("mirror", "periodic", "asym mirror"):
            SMALL_CASES.append((f"bounce_{w}x{h}_{bc.replace(' ', '_')}", (h, w), "float32", [3.0, 2.0], [0, 1],
                                [bc], None))
# uint8 planes with radius <= 4 and 4-sample-aligned rows (the register-streaming
# kernel's domain): every boundary condition, partial strips, tiny planes, stacks
for bc in ("mirror", "asym mirror", "periodic", "asym periodic", "add zeros", "add max", "add min", "zero order"):
    SMALL_CASES.append((f"u8_64x40_s1_o01_{bc.replace(' ', '_')}", (40, 64), "uint8", [1.0], [0, 1], [bc], None))
SMALL_CASES.append(("u8_260x21_s1_o20_mirror", (21, 260), "uint8", [1.0], [2, 0], ["mirror"], None))
SMALL_CASES.append(("u8_12x9_s1_o11_periodic", (9, 12), "uint8", [1.0], [1, 1], ["periodic"], None))
SMALL_CASES.append(("u8_8x5_s1_o00_mirror", (5, 8), "uint8", [1.0], [0, 0], ["mirror"], None))
SMALL_CASES.append(("u8_stack3_64x40_s1", (3, 40, 64), "uint8", [1.0, 1.2, 0.0], [0, 0, 0], ["mirror"], None))
SMALL_CASES.append(("u8_520x70_s07_o00_asym_mirror", (70, 520), "uint8", [0.7, 1.3], [0, 0], ["asym mirror"],
                    None))
# width-1 lines (dim skipped) and single pixels
SMALL_CASES.append(("f32_1x40_s2", (40, 1), "float32", [2.0], [0], ["mirror"], None))
SMALL_CASES.append(("f32_1x1_s2", (1, 1), "float32", [2.0], [0], ["mirror"], None))

# Boundary conditions whose extension is computed from the edge values rather than
# copied (copy_buffer.cpp:380-513, 551-589): own fixture file (extrap_cases.npz),
# same tuple layout as SMALL_CASES.  Widths below the radius make the antisymmetric
# reflection bounce; width 1 takes the reference's fallbacks (constant, first order).
VALUE_BCS = ("antisym reflect", "first order", "second order", "third order")
EXTRAP_CASES = []
for bc in VALUE_BCS:
    t = bc.replace(" ", "_")
    EXTRAP_CASES += [
        (f"f32_64x48_s2_o00_{t}", (48, 64), "float32", [2.0], [0, 0], [bc], None),
        (f"u8_61x37_s3_o10_{t}", (37, 61), "uint8", [3.0], [1, 0], [bc], None),
        (f"u8_50x40_s2_o00_{t}_u8out", (40, 50), "uint8", [2.0], [0, 0], [bc], "uint8"),
        (f"f64_33x20_s2_o02_{t}", (20, 33), "float64", [2.0], [0, 2], [bc], None),
        (f"f32_37x29x11_s221_{t}", (11, 29, 37), "float32", [2.0, 2.0, 1.0], [0], [bc], None),
        (f"f32_130x70_s8_o00_{t}", (70, 130), "float32", [8.0], [0, 0], [bc], None),
    ]
    for w in (1, 2, 3, 5):
        EXTRAP_CASES.append((f"bounce_{w}x9_{t}", (9, w), "float32", [3.0, 2.0], [0, 1], [bc], None))
EXTRAP_CASES.append(("f32_40x30_first_order_x_mirror_y", (30, 40), "float32", [2.0], [0, 1],
                     ["first order", "mirror"], None))
EXTRAP_CASES.append(("f32_40x30_periodic_x_third_order_y", (30, 40), "float32", [2.0], [1, 0],
                     ["periodic", "third order"], None))
EXTRAP_CASES.append(("s16_45x33_s2_antisym_reflect", (33, 45), "int16", [2.0], [0, 0], ["antisym reflect"], None))

# The derivative family (dip::Gradient, GradientMagnitude, Hessian, Laplace with the FIR
# method): (name, function, numpy shape, dtype, sigmas (dip order), boundary list).
DERIV_CASES = []
for fn in ("gradient", "gradmag", "hessian", "laplace"):
    DERIV_CASES += [
        (f"{fn}_f32_64x48_s2", fn, (48, 64), "float32", [2.0], ["mirror"]),
        (f"{fn}_u8_97x61_s15_periodic", fn, (61, 97), "uint8", [1.5], ["periodic"]),
        (f"{fn}_f64_33x20_s3", fn, (20, 33), "float64", [3.0], ["mirror"]),
        (f"{fn}_f32_37x29x11_s221", fn, (11, 29, 37), "float32", [2.0, 2.0, 1.0], ["mirror"]),
        (f"{fn}_f32_1x40_s2", fn, (40, 1), "float32", [2.0], ["mirror"]),
        (f"{fn}_f32_50x40_s0x3", fn, (40, 50), "float32", [0.0, 3.0], ["asym mirror"]),
    ]

# 2-D uint8 images whose Gaussian derivatives all have radius <= 4 (sigma <= 1): the
# one-launch form of the family (several strips, a partial last strip, every index-map
# boundary condition).
for fn in ("gradient", "gradmag", "hessian", "laplace"):
    DERIV_CASES += [
        (f"{fn}_u8_96x64_s1", fn, (64, 96), "uint8", [1.0], ["mirror"]),
        (f"{fn}_u8_128x40_s08_periodic", fn, (40, 128), "uint8", [0.8], ["periodic"]),
        (f"{fn}_u8_64x48_s1x07_mixedbc", fn, (48, 64), "uint8", [1.0, 0.7], ["asym mirror", "add zeros"]),
        (f"{fn}_u8_300x20_s1_zero_order", fn, (20, 300), "uint8", [1.0], ["zero order"]),
        (f"{fn}_u8_64x32_s1_addmax", fn, (32, 64), "uint8", [1.0], ["add max", "asym periodic"]),
    ]

# The Dgg family and the structure tensor ("structure:<tensor sigma>"), same layout.
DERIV2_CASES = []
for fn in ("dgg", "laplace+dgg", "laplace-dgg", "structure:3"):
    t = fn.replace("+", "p").replace("-", "m").replace(":", "")
    DERIV2_CASES += [
        (f"{t}_f32_64x48_s2", fn, (48, 64), "float32", [2.0], ["mirror"]),
        (f"{t}_u8_97x61_s15_periodic", fn, (61, 97), "uint8", [1.5], ["periodic"]),
        (f"{t}_f64_33x20_s3", fn, (20, 33), "float64", [3.0], ["mirror"]),
        (f"{t}_f32_37x29x11_s221", fn, (11, 29, 37), "float32", [2.0, 2.0, 1.0], ["mirror"]),
        (f"{t}_f32_1x40_s2", fn, (40, 1), "float32", [2.0], ["mirror"]),
    ]
# a flat image: the gradient is zero everywhere and SafeDivide gives 0
DERIV2_CASES.append(("dgg_flat_f32_30x20", "dgg", (20, 30), "flat", [2.0], ["mirror"]))
# the one-launch form (2-D uint8, sigma <= 1)
for fn in ("dgg", "laplace+dgg", "laplace-dgg"):
    t = fn.replace("+", "p").replace("-", "m")
    DERIV2_CASES += [
        (f"{t}_u8_96x64_s1", fn, (64, 96), "uint8", [1.0], ["mirror"]),
        (f"{t}_u8_300x40_s08_periodic", fn, (40, 300), "uint8", [0.8], ["periodic"]),
        (f"{t}_u8_64x48_s1x07_mixedbc", fn, (48, 64), "uint8", [1.0, 0.7], ["asym mirror", "add zeros"]),
    ]

TAP_SIGMAS = [0.5, 0.8, 1.0, 1.5, 2.0, 3.0, 4.0, 5.0, 8.0, 16.0, 31.7]

# Full-size workloads: (name, numpy shape, dtype, seed, impulses, sigmas, orders, bc, out dtype)
LARGE_CASES = []
for s in (1.0, 4.0, 16.0):
    for o in ((0, 0), (1, 0), (0, 2)):
        LARGE_CASES.append((f"u8_8192sq_s{int(s)}_o{o[0]}{o[1]}", (8192, 8192), "uint8", 2, 0.0, [s], list(o),
                            ["mirror"], None))
for bc in ("mirror", "periodic"):
    LARGE_CASES.append((f"f32_512x512x256_s221_{bc}", (256, 512, 512), "float32", 3, 0.01, [2.0, 2.0, 1.0], [0],
                        [bc], None))
LARGE_CASES.append(("f64_256x256x128_s221_mirror", (128, 256, 256), "float64", 4, 0.01, [2.0, 2.0, 1.0], [0],
                    ["mirror"], None))

# dip::GaussIIR: (name, numpy shape, dtype, sigmas (dip order), orders, boundary list,
# filter orders (() = default), design method, truncation, out dtype or None)
DTF, FB = "discrete time fit", "forward backward"
IIR_CASES = [
    ("iir_f32_64x48_s12", (48, 64), "float32", [12.0], [0, 0], ["mirror"], (), DTF, 3.0, None),
    ("iir_u8_97x61_s11_o10_periodic", (61, 97), "uint8", [11.0], [1, 0], ["periodic"], (), DTF, 3.0, None),
    ("iir_f64_80x40_s15_o02", (40, 80), "float64", [15.0], [0, 2], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_37x29x11_s11", (11, 29, 37), "float32", [11.0, 11.0, 12.0], [0], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_50x40_o30", (40, 50), "float32", [4.0], [3, 0], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_50x40_o04", (40, 50), "float32", [4.0], [0, 4], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_50x40_s1", (40, 50), "float32", [1.0], [0, 1], ["mirror"], (), DTF, 3.0, None),
    ("iir_u8_50x40_u8out", (40, 50), "uint8", [3.0], [0, 0], ["mirror"], (), DTF, 3.0, "uint8"),
    ("iir_s16_45x33", (33, 45), "int16", [2.5], [0, 0], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_1x40", (40, 1), "float32", [12.0], [0], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_40x30_s0x5", (30, 40), "float32", [0.0, 5.0], [0, 1], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_2x3_tiny", (3, 2), "float32", [2.0], [0, 1], ["mirror"], (), DTF, 3.0, None),
    ("iir_f32_60x50_t2", (50, 60), "float32", [6.0], [1, 1], ["mirror"], (), DTF, 2.0, None),
    ("iir_f32_60x50_t5", (50, 60), "float32", [6.0], [2, 0], ["mirror"], (), DTF, 5.0, None),
]
for o in range(5):
    for fo in (1, 2, 3, 4, 5, 6):
        IIR_CASES.append((f"iir_dtf_o{o}_fo{fo}", (20, 40), "float32", [5.0, 3.0], [o, 0], ["mirror"], (fo,), DTF,
                          3.0, None))
    for fo in (3, 4, 5, 2):
        IIR_CASES.append((f"iir_fb_o{o}_fo{fo}", (20, 40), "float64", [5.0, 3.0], [0, o], ["mirror"], (fo,), FB,
                          3.0, None))
for bc in ("asym mirror", "periodic", "asym periodic", "add zeros", "zero order", "antisym reflect",
           "first order", "second order", "third order"):
    t = bc.replace(" ", "_")
    IIR_CASES.append((f"iir_f32_40x20_{t}", (20, 40), "float32", [6.0, 4.0], [1, 2], [bc], (), DTF, 3.0, None))
    IIR_CASES.append((f"iir_line_3_{t}", (1, 3), "float64", [2.0], [0, 0], [bc], (), DTF, 3.0, None))

# The derivative family and the structure tensor with the IIR method ("best" with a sigma
# above 10, or "gaussiir"): (name, function, numpy shape, dtype, sigmas, boundary list, method)
DERIV_IIR_CASES = []
for fn in ("gradient", "gradmag", "hessian", "laplace", "dgg", "laplace-dgg", "structure:12", "structure:3"):
    t = fn.replace("+", "p").replace("-", "m").replace(":", "")
    DERIV_IIR_CASES += [
        (f"{t}_f32_64x48_s12_best", fn, (48, 64), "float32", [12.0], ["mirror"], "best"),
        (f"{t}_u8_61x37_s3_iir", fn, (37, 61), "uint8", [3.0], ["periodic"], "gaussiir"),
        (f"{t}_f64_40x30_s11_best", fn, (30, 40), "float64", [11.0], ["mirror"], "best"),
    ]


def iir_flags(fn, sigmas, method):
    """(derivatives by IIR, structure-tensor smoothing by IIR) as dip::Gauss resolves
    `method` (derivative.cpp:35-58)."""
    def one(sig):
        m = method[5:] if method.startswith("gauss") and method != "gauss" else method
        if m == "best":
            return any(s > 10 for s in sig)
        return m == "iir"
    tsig = [float(fn.split(":")[1])] if fn.startswith("structure:") else sigmas
    return one(sigmas), one(tsig)

# dip::GaussFT (spatial in and out): (name, numpy shape, dtype, sigmas, orders, boundary list
# (empty: no extension), truncation).  Compared with a tolerance: the FFTs round differently.
FT_CASES = [
    ("ft_f32_64x48_s05", (48, 64), "float32", [0.5], [0, 0], ["mirror"], 3.0),
    ("ft_u8_61x37_s06_o10_periodic", (37, 61), "uint8", [0.6], [1, 0], ["periodic"], 3.0),
    ("ft_f64_40x30_s2_o40", (30, 40), "float64", [2.0], [4, 0], ["mirror"], 3.0),
    ("ft_f32_40x29x11_s07", (11, 29, 40), "float32", [0.7, 0.7, 0.5], [0], ["mirror"], 3.0),
    ("ft_f32_50x40_s0_o11", (40, 50), "float32", [0.0], [1, 1], ["mirror"], 3.0),
    ("ft_f32_64x48_s3_noext", (48, 64), "float32", [3.0], [0, 0], [], 3.0),
    ("ft_f32_97x61_s15_o21", (61, 97), "float32", [1.5], [2, 1], ["mirror"], 3.0),
    ("ft_f32_46x30_add_zeros", (30, 46), "float32", [0.5], [0, 1], ["add zeros"], 3.0),
    ("ft_f32_40x30_asym_mirror", (30, 40), "float32", [0.7], [1, 0], ["asym mirror"], 3.0),
    ("ft_f32_40x30_zero_order", (30, 40), "float32", [2.5], [0, 2], ["zero order"], 3.0),
    ("ft_f32_40x30_asym_periodic", (30, 40), "float32", [1.2], [3, 0], ["asym periodic"], 3.0),
    ("ft_f32_1x42", (42, 1), "float32", [0.5], [0], ["mirror"], 3.0),
    ("ft_f64_33x20_s04_o50", (20, 33), "float64", [0.4], [5, 0], ["mirror"], 3.0),
    ("ft_f32_45x33_mixed_bc", (33, 45), "float32", [0.6, 1.0], [0, 1], ["periodic", "add zeros"], 3.0),
    ("ft_f32_64x64_t2", (64, 64), "float32", [1.0], [1, 0], ["mirror"], 2.0),
    ("ft_s16_45x33", (33, 45), "int16", [0.5], [0, 0], ["mirror"], 3.0),
]
# Odd padded length along the real-to-complex dim: the reference's output is not shift
# invariant there (its odd-length real DFT path), so these are checked against the
# oracle's exact convolution and the shift invariance of the result, not the reference.
FT_ODD_CASES = [
    ("ftodd_f32_37x29x11", (11, 29, 37), "float32", [0.7, 0.7, 0.5], [0], ["mirror"], 3.0),
    ("ftodd_f32_40x30_add_zeros", (30, 40), "float32", [0.5], [0, 1], ["add zeros"], 3.0),
    ("ftodd_f32_41_noext", (41,), "float32", [0.5], [0], [], 3.0),
    ("ftodd_f32_45x5_noext_o1", (5, 45), "float32", [0.8], [1, 0], [], 3.0),
]

# The derivative family and the structure tensor through the FT method ("best" with a sigma
# below 0.8, or "gaussft"); compared with a tolerance (FFT rounding).  Dgg is left out: its
# division by the squared gradient magnitude amplifies rounding where the gradient vanishes.
# dip::GaussFT with a frequency-domain input or output (no boundary condition):
# (name, numpy shape, dtype, sigmas (dip order), orders, inRepresentation, outRepresentation).
# Real inputs are 1-D or 2-D with an even x length: the reference's real-to-complex path
# gives no exact DFT for an odd length of that dim, nor at the Nyquist plane of a third
# dim (see the oracle tests); complex inputs take its exact complex path.
FT_REPR_CASES = [
    ("ftr_f32_64x48_s2_spec", (48, 64), "float32", [2.0], [0], "spatial", "frequency"),
    ("ftr_u8_62x40_s15_o10_spec", (40, 62), "uint8", [1.5], [1, 0], "spatial", "frequency"),
    ("ftr_f64_40x31_s3_o02_spec", (31, 40), "float64", [3.0], [0, 2], "spatial", "frequency"),
    ("ftr_c64_20x16x10_s1_spec", (10, 16, 20), "complex64", [1.0, 1.0, 0.8], [0], "spatial", "frequency"),
    ("ftr_f32_1d_50_s05_o3_spec", (50,), "float32", [0.5], [3], "spatial", "frequency"),
    ("ftr_c64_45x33_s2_spec", (33, 45), "complex64", [2.0], [0, 1], "spatial", "frequency"),
    ("ftr_c64_64x48_s2_inverse", (48, 64), "complex64", [2.0], [0], "frequency", "spatial"),
    ("ftr_c128_33x20_o11_inverse", (20, 33), "complex128", [1.0], [1, 1], "frequency", "spatial"),
    ("ftr_c64_40x30_s1_spec_to_spec", (30, 40), "complex64", [1.0, 2.0], [0], "frequency", "frequency"),
    ("ftr_c64_37x29x5_inverse", (5, 29, 37), "complex64", [1.5, 0.0, 1.0], [0], "frequency", "spatial"),
]

DERIV_FT_CASES = []
for fn in ("gradient", "gradmag", "hessian", "laplace", "structure:0.6", "structure:3"):
    t = fn.replace(":", "").replace(".", "")
    DERIV_FT_CASES += [
        (f"{t}_f32_64x48_s05_best", fn, (48, 64), "float32", [0.5], ["mirror"], "best"),
        (f"{t}_u8_62x38_s2_ft", fn, (38, 62), "uint8", [2.0], ["periodic"], "gaussft"),
        (f"{t}_f64_42x30_s07_best", fn, (30, 42), "float64", [0.7], ["mirror"], "best"),
    ]


def method_flags(fn, sigmas, method):
    """Which Gaussian the derivatives (first) and the structure tensor's smoothing
    (second) run, "fir" / "iir" / "ft", as dip::Gauss resolves `method`."""
    def one(sig):
        m = "best" if method == "gauss" else (method[5:] if method.startswith("gauss") else method)
        if m == "best":
            if any(0 < s < 0.8 for s in sig):
                return "ft"
            return "iir" if any(s > 10 for s in sig) else "fir"
        return m
    tsig = [float(fn.split(":")[1])] if fn.startswith("structure:") else sigmas
    return one(sigmas), one(tsig)


# dip::SeparableConvolution with user filters (convolution.cpp:67-162, 202-297):
# (name, numpy shape, dtype, filters, boundary list, process or None, out dtype or None),
# each filter a (taps in the user's order, symmetry string, origin) triple; one filter is
# used for every dim, else one per dim (dip order).  Taps are seeded N(0,1) values;
# "neg" filters have only negative taps, so uint8 zeros give -0.0 products (the general
# and d-* forms start their sums from +0, the others from the centre product).
def sep_taps(n, seed, neg=False):
    t = np.random.default_rng(seed).standard_normal(n)
    return list(-np.abs(t)) if neg else list(t)


SEP_SYMS = ("general", "even", "odd", "conj", "d-even", "d-odd", "d-conj")
SEP_CASES = []
# every symmetry, odd and even stored lengths, centre origin, both passes
for i, sym in enumerate(SEP_SYMS):
    t = sym.replace("-", "")
    for n in (3, 4):
        SEP_CASES.append((f"sep_{t}_n{n}_f32_45x33", (33, 45), "float32",
                          [(sep_taps(n, 100 + 10 * i + n), sym, -1)], ["mirror"], None, None))
    SEP_CASES.append((f"sep_{t}_n1_f32_30x20", (20, 30), "float32", [(sep_taps(1, 200 + i), sym, -1)],
                      ["mirror"], None, None))
# origins: 0, the last index, off-centre; per-dim filters of different symmetries
for i, sym in enumerate(SEP_SYMS):
    t = sym.replace("-", "")
    n = 4
    size = n if sym == "general" else (2 * n - 1 if sym in ("even", "odd", "conj") else 2 * n)
    for org in (0, size - 1, 1):
        SEP_CASES.append((f"sep_{t}_origin{org}_f32_40x31", (31, 40), "float32",
                          [(sep_taps(n, 300 + 10 * i + org), sym, org), (sep_taps(3, 400 + i), "even", org % 5)],
                          ["mirror", "periodic"], None, None))
# every index-map boundary condition with lines shorter than the filter (bounce, wrap)
for j, bc in enumerate(("mirror", "asym mirror", "periodic", "asym periodic", "add zeros", "add max", "add min",
                        "zero order")):
    t = bc.replace(" ", "_")
    for i, sym in enumerate(("general", "d-odd", "even")):
        SEP_CASES.append((f"sep_{sym.replace('-', '')}_{t}_short_5x3", (3, 5), "float32",
                          [(sep_taps(7, 500 + 10 * j + i), sym, 2 if sym == "general" else -1)], [bc], None, None))
    SEP_CASES.append((f"sep_general_{t}_f32_37x29", (29, 37), "float32",
                      [(sep_taps(5, 600 + j), "general", 1), (sep_taps(4, 620 + j), "d-even", 0)], [bc], None, None))
# conditions computed from the edge values
for j, bc in enumerate(("antisym reflect", "first order", "second order", "third order")):
    t = bc.replace(" ", "_")
    SEP_CASES.append((f"sep_general_{t}_f32_37x29", (29, 37), "float32",
                      [(sep_taps(6, 700 + j), "general", 4), (sep_taps(3, 720 + j), "odd", -1)], [bc], None, None))
    SEP_CASES.append((f"sep_deven_{t}_short_4x3", (3, 4), "float32",
                      [(sep_taps(5, 740 + j), "d-even", 1)], [bc], None, None))
# process masks, meaningless and empty filters
SEP_CASES += [
    ("sep_process_x_f32_40x30", (30, 40), "float32", [(sep_taps(5, 800), "general", -1)], ["mirror"], [1, 0], None),
    ("sep_process_y_f32_40x30", (30, 40), "float32", [(sep_taps(3, 801), "odd", -1)], ["mirror"], [0, 1], None),
    ("sep_identity_x_f32_40x30", (30, 40), "float32", [([1.0], "general", -1), (sep_taps(3, 802), "even", -1)],
     ["mirror"], None, None),
    ("sep_empty_y_f32_40x30", (30, 40), "float32", [(sep_taps(4, 803), "d-odd", -1), ([], "general", -1)],
     ["mirror"], None, None),
    ("sep_process_none_u8_40x30", (30, 40), "uint8", [(sep_taps(3, 804), "even", -1)], ["mirror"], [0, 0], None),
]
# signed zeros: uint8 zeros times negative taps
for i, sym in enumerate(("general", "d-even", "d-odd", "even", "odd")):
    SEP_CASES.append((f"sep_{sym.replace('-', '')}_neg_u8_48x40", (40, 48), "uint8",
                      [(sep_taps(3, 850 + i, neg=True), sym, -1)], ["add zeros"], None, None))
# other types, dims and lengths: uint8 and dfloat input, 3-D, 1-D, long filters (past
# the fused kernels' reach), narrow output types
SEP_CASES += [
    ("sep_even_u8_64x48", (48, 64), "uint8", [(sep_taps(5, 900), "even", -1)], ["mirror"], None, None),
    ("sep_odd_f64_45x33", (33, 45), "float64", [(sep_taps(4, 901), "odd", -1)], ["periodic"], None, None),
    ("sep_general_f64_45x33_origin0", (33, 45), "float64", [(sep_taps(6, 902), "general", 0)], ["mirror"], None,
     None),
    ("sep_deven_s16_45x33", (33, 45), "int16", [(sep_taps(3, 903), "d-even", -1)], ["mirror"], None, None),
    ("sep_mixed_f32_21x17x9", (9, 17, 21), "float32",
     [(sep_taps(3, 904), "even", -1), (sep_taps(4, 905), "general", 1), (sep_taps(2, 906), "d-odd", 0)],
     ["mirror", "periodic", "add zeros"], None, None),
    ("sep_general_1d_f32_200", (200,), "float32", [(sep_taps(9, 907), "general", 6)], ["mirror"], None, None),
    ("sep_even_r40_f32_130x90", (90, 130), "float32", [(sep_taps(41, 908), "even", -1)], ["mirror"], None, None),
    ("sep_general_n61_f32_150x80", (80, 150), "float32", [(sep_taps(61, 909), "general", 20)], ["periodic"], None,
     None),
    ("sep_dodd_n30_f32_120x70", (70, 120), "float32", [(sep_taps(30, 910), "d-odd", -1)], ["mirror"], None, None),
    ("sep_even_f32_out_u8_64x48", (48, 64), "float32", [([0.25, 0.5], "even", -1)], ["mirror"], None, "uint8"),
    ("sep_odd_u8_out_s16_64x48", (48, 64), "uint8", [([-1.0, 0.0], "odd", -1)], ["mirror"], None, "int16"),
    ("sep_even_u8_r4_256x64", (64, 256), "uint8", [(sep_taps(5, 911), "even", -1)], ["periodic"], None, None),
]


# Complex separable convolution (dip::SeparableConvolution with a complex buffer type):
# (name, numpy shape, dtype, [(taps, symmetry, origin, is_complex), ...], boundary list,
# process or None, out dtype or None).  Complex taps are Python complex values.
def csep_taps(n, seed):
    rng = np.random.default_rng(seed)
    return list(rng.standard_normal(n) + 1j * rng.standard_normal(n))


CSEP_CASES = []
# complex filters of every symmetry (conj / d-conj act as such only here), real input
for i, sym in enumerate(SEP_SYMS):
    t = sym.replace("-", "")
    for n in (3, 4):
        CSEP_CASES.append((f"csep_{t}_n{n}_f32_37x29", (29, 37), "float32",
                           [(csep_taps(n, 1100 + 10 * i + n), sym, -1, True)], ["mirror"], None, None))
# complex filters on complex input, off-centre origins (d-conj origin 0: one sample
# past the border), per-dim filters
for i, sym in enumerate(SEP_SYMS):
    t = sym.replace("-", "")
    CSEP_CASES.append((f"csep_{t}_c64_origin0_31x23", (23, 31), "complex64",
                       [(csep_taps(4, 1200 + i), sym, 0, True), (csep_taps(3, 1220 + i), "general", 2, True)],
                       ["mirror", "periodic"], None, None))
# real filters of a complex filter array (stored as (f, 0)) next to a complex one;
# real filters with conj symmetry act as even when no filter of the array is complex
CSEP_CASES += [
    ("csep_mixed_real_complex_f32_40x30", (30, 40), "float32",
     [(sep_taps(5, 1300), "even", -1, False), (csep_taps(3, 1301), "conj", -1, True)], ["mirror"], None, None),
    ("csep_mixed_dconj_real_u8_40x30", (30, 40), "uint8",
     [(sep_taps(3, 1302, neg=True), "d-conj", -1, False), (csep_taps(2, 1303), "odd", -1, True)],
     ["add zeros"], None, None),
    ("csep_real_filters_c64_45x33", (33, 45), "complex64",
     [(sep_taps(5, 1304), "conj", -1, False), (sep_taps(4, 1305), "d-odd", 1, False)], ["mirror"], None, None),
    ("csep_real_general_c128_45x33", (33, 45), "complex128",
     [(sep_taps(6, 1306), "general", 2, False)], ["periodic"], None, None),
]
# every boundary condition, including those computed from the edge values (applied to
# the real and imaginary parts separately), with short lines
for j, bc in enumerate(("mirror", "asym mirror", "periodic", "asym periodic", "add zeros", "add max", "add min",
                        "zero order", "antisym reflect", "first order", "second order", "third order")):
    t = bc.replace(" ", "_")
    CSEP_CASES.append((f"csep_general_{t}_c64_33x25", (25, 33), "complex64",
                       [(csep_taps(5, 1400 + j), "general", 1, True), (csep_taps(4, 1420 + j), "d-conj", -1, True)],
                       [bc], None, None))
    CSEP_CASES.append((f"csep_even_{t}_short_5x3", (3, 5), "float32",
                       [(csep_taps(6, 1440 + j), "even", -1, True)], [bc], None, None))
# double precision buffers, 3-D, 1-D, process masks, long filters, narrower output
CSEP_CASES += [
    ("csep_conj_f64_45x33", (33, 45), "float64", [(csep_taps(4, 1500), "conj", -1, True)], ["mirror"], None, None),
    ("csep_dconj_s32_45x33", (33, 45), "int32", [(csep_taps(3, 1501), "d-conj", 0, True)], ["periodic"], None,
     None),
    ("csep_mixed_c64_17x13x9", (9, 13, 17), "complex64",
     [(csep_taps(3, 1502), "conj", -1, True), (sep_taps(4, 1503), "general", 1, False),
      (csep_taps(2, 1504), "d-odd", 3, True)], ["mirror", "periodic", "add zeros"], None, None),
    ("csep_general_1d_c128_150", (150,), "complex128", [(csep_taps(9, 1505), "general", 6, True)], ["mirror"], None,
     None),
    ("csep_process_y_u16_40x30", (30, 40), "uint16", [(csep_taps(3, 1506), "odd", -1, True)], ["mirror"], [0, 1],
     None),
    ("csep_identity_c64_40x30", (30, 40), "complex64", [([1.0], "general", -1, False)], ["mirror"], None, None),
    ("csep_long_conj_f32_120x70", (70, 120), "float32", [(csep_taps(25, 1507), "conj", -1, True)], ["mirror"], None,
     None),
    ("csep_c128_out_c64_40x30", (30, 40), "complex128", [(csep_taps(4, 1508), "even", -1, True)], ["mirror"], None,
     "complex64"),
]


# The masked forms: dip::NormalizedConvolution ("nc"), NormalizedDifferentialConvolution
# ("ndc:<dim>") and the masked StructureTensor ("structure:<tensor sigma>").
# (name, function, numpy shape, dtype, mask ("bin": seeded 70 % true, "f32": seeded
# uniform [0, 1) with zeros), sigmas (dip order), boundary list)
MASKED_CASES = []
for fn in ("nc", "ndc:0", "ndc:1", "structure:3"):
    t = fn.replace(":", "")
    MASKED_CASES += [
        (f"{t}_f32_52x40_bin", fn, (40, 52), "float32", "bin", [2.0], ["mirror"]),
        (f"{t}_u8_45x33_f32mask", fn, (33, 45), "uint8", "f32", [1.5], ["periodic"]),
        (f"{t}_f64_30x24_bin", fn, (24, 30), "float64", "bin", [1.0, 2.0], ["mirror"]),
    ]
MASKED_CASES.append(("ndc2_f32_20x16x9_bin", "ndc:2", (9, 16, 20), "float32", "bin", [1.0], ["mirror"]))
MASKED_CASES.append(("structure2_f32_20x16x9_bin", "structure:2", (9, 16, 20), "float32", "bin", [1.0], ["mirror"]))


def make_mask(shape, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "bin":
        return rng.random(shape) < 0.7
    m = rng.random(shape).astype(np.float32)
    m[rng.random(shape) < 0.2] = 0.0  # zeros: SafeDivide's 0 where the mask's Gaussian vanishes locally
    return m
