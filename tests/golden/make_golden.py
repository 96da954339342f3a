"""Writes the golden fixtures under tests/golden/ from the REFERENCE library itself
(oracle/_ref/libdipref.so, built from the reference sources by oracle/ref.mk).

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py [--only extrap|deriv|iir|ft]

Outputs (all small):
  extrap_cases.npz   input and dip::Gauss output of tests/cases.py:EXTRAP_CASES (the
                     boundary conditions that extrapolate from the edge values)
  deriv_cases.npz    input and result (tensor elements on the first axis) of
                     tests/cases.py:DERIV_CASES (dip::Gradient, GradientMagnitude,
                     Hessian, Laplace with the FIR method)
  deriv2_cases.npz   the same for tests/cases.py:DERIV2_CASES (dip::Dgg, LaplacePlusDgg,
                     LaplaceMinusDgg, StructureTensor)
  iir_cases.npz      input and dip::GaussIIR output of tests/cases.py:IIR_CASES
  deriv_iir_cases.npz  the derivative family and structure tensor with the IIR method
                     (tests/cases.py:DERIV_IIR_CASES)
  ft_cases.npz       input and dip::GaussFT output of tests/cases.py:FT_CASES
  deriv_ft_cases.npz the derivative family and structure tensor with the FT method
                     (tests/cases.py:DERIV_FT_CASES)
  sep_cases.npz      input and dip::SeparableConvolution output of tests/cases.py:SEP_CASES
                     (user filters of every symmetry, origins, process masks, every
                     boundary condition)
  taps.npz           MakeHalfGaussian(sigma, order, 3, sfloat/dfloat) for TAP_SIGMAS x orders 0..3
  small_cases.npz    input and dip::Gauss output of every case in tests/cases.py:SMALL_CASES
  large_cases.json   SHA-256 of the exact output bytes + sampled pixels for LARGE_CASES
                     (inputs regenerated from their seeds by tests/cases.py:make_input)
  jacobian.npz       J^T w for dip::Gauss on small images, J assembled column by column
                     from the reference's impulse responses (dfloat), w ~ N(0,1) seeded
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402
import oracle_ctypes as oc  # noqa: E402


def sample_idx(n, k=64, seed=7):
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([np.arange(min(n, 8)), n - 1 - np.arange(min(n, 8)), rng.integers(0, n, k)]))


def write_cases(case_list, fname, seed0):
    arrs = {}
    for i, (name, shape, dt, sig, ords, bc, odt) in enumerate(case_list):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        y = oc.gauss_fir(x, sig, ords, bc, 3.0, out_dtype=odt, nthreads=1, use_ref=True)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = y
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_derivative_cases(case_list=None, fname="deriv_cases.npz", seed0=4000):
    arrs = {}
    for i, (name, fn, shape, dt, sig, bc) in enumerate(case_list or cases.DERIV_CASES):
        x = np.ones(shape, np.float32) if dt == "flat" else cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.ref_derivative(fn, x, sig, bc)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_iir_cases(fname="iir_cases.npz", seed0=6000):
    arrs = {}
    for i, (name, shape, dt, sig, ords, bc, fo, method, trunc, odt) in enumerate(cases.IIR_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.gauss_iir(x, sig, ords, bc, fo, method, trunc, out_dtype=odt, nthreads=1,
                                            use_ref=True)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_derivative_iir_cases(fname="deriv_iir_cases.npz", seed0=8000):
    arrs = {}
    for i, (name, fn, shape, dt, sig, bc, method) in enumerate(cases.DERIV_IIR_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.ref_derivative(fn, x, sig, bc, method)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_ft_cases(fname="ft_cases.npz", seed0=9500):
    arrs = {}
    for i, (name, shape, dt, sig, ords, bc, trunc) in enumerate(cases.FT_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.ref_gauss_ft(x, sig, ords, bc, trunc)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_ft_repr_cases(fname="ft_repr_cases.npz", seed0=9800):
    arrs = {}
    for i, (name, shape, dt, sig, ords, ir, orr) in enumerate(cases.FT_REPR_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.ref_gauss_ft_repr(x, sig, ords, 3.0, ir, orr)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_masked_cases(fname="masked_cases.npz", seed0=13000):
    arrs = {}
    for i, (name, fn, shape, dt, mk, sig, bc) in enumerate(cases.MASKED_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        m = cases.make_mask(shape, mk, seed=seed0 + 500 + i)
        arrs[name + "__in"] = x
        arrs[name + "__mask"] = m
        arrs[name + "__out"] = oc.ref_derivative_masked(fn, x, m, sig, bc)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_derivative_ft_cases(fname="deriv_ft_cases.npz", seed0=9900):
    arrs = {}
    for i, (name, fn, shape, dt, sig, bc, method) in enumerate(cases.DERIV_FT_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.ref_derivative(fn, x, sig, bc, method)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_sep_cases(fname="sep_cases.npz", seed0=11000):
    arrs = {}
    for i, (name, shape, dt, filters, bc, proc, odt) in enumerate(cases.SEP_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        taps = [f[0] for f in filters]
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.separable(x, taps, [f[1] for f in filters], bc, proc, odt,
                                            origins=[f[2] for f in filters], use_ref=True)
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def write_csep_cases(fname="csep_cases.npz", seed0=12000):
    """Complex separable convolution fixtures (complex filters and/or complex images),
    written by the reference."""
    arrs = {}
    for i, (name, shape, dt, filters, bc, proc, odt) in enumerate(cases.CSEP_CASES):
        x = cases.make_input(shape, dt, seed=seed0 + i)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = oc.separable(x, [f[0] for f in filters], [f[1] for f in filters], bc, proc, odt,
                                            origins=[f[2] for f in filters], use_ref=True,
                                            is_complex=[f[3] for f in filters])
    np.savez_compressed(os.path.join(HERE, fname), **arrs)


def main():
    assert oc.ref_available(), "build oracle/_ref/libdipref.so first (make -f oracle/ref.mk)"
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    if only in (None, "sep"):
        write_sep_cases()
    if only in (None, "csep"):
        write_csep_cases()
    if only in (None, "extrap"):
        write_cases(cases.EXTRAP_CASES, "extrap_cases.npz", 3000)
    if only in (None, "ft"):
        write_ft_cases()
        write_ft_repr_cases()
        write_derivative_ft_cases()
    if only in (None, "iir"):
        write_iir_cases()
        write_derivative_iir_cases()
    if only in (None, "deriv"):
        write_derivative_cases()
        write_derivative_cases(cases.DERIV2_CASES, "deriv2_cases.npz", 5000)
        write_masked_cases()
    if only is not None:
        return
    # taps
    taps = {}
    for s in cases.TAP_SIGMAS:
        for o in range(4):
            for dd in (0, 1):
                taps[f"s{s}_o{o}_d{dd}"] = oc.half_gaussian(s, o, 3.0, bool(dd), use_ref=True)
    np.savez_compressed(os.path.join(HERE, "taps.npz"), **taps)

    # small full-output cases
    arrs = {}
    for i, (name, shape, dt, sig, ords, bc, odt) in enumerate(cases.SMALL_CASES):
        x = cases.make_input(shape, dt, seed=1000 + i)
        y = oc.gauss_fir(x, sig, ords, bc, 3.0, out_dtype=odt, nthreads=1, use_ref=True)
        arrs[name + "__in"] = x
        arrs[name + "__out"] = y
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **arrs)

    # large cases: hashes of the exact output bytes + sampled pixels
    large = {}
    for (name, shape, dt, seed, imp, sig, ords, bc, odt) in cases.LARGE_CASES:
        x = cases.make_input(shape, dt, seed, imp)
        y = oc.gauss_fir(x, sig, ords, bc, 3.0, out_dtype=odt, nthreads=8, use_ref=True)
        flat = y.reshape(-1)
        idx = sample_idx(flat.size)
        large[name] = {
            "input_sha256": hashlib.sha256(x.tobytes()).hexdigest(),
            "output_sha256": hashlib.sha256(y.tobytes()).hexdigest(),
            "output_dtype": str(y.dtype),
            "sample_index": idx.tolist(),
            "sample_value": [float(v) for v in flat[idx]],
            "sum": float(flat.astype(np.float64).sum()),
        }
        print(name, large[name]["output_sha256"][:16], flush=True)
    with open(os.path.join(HERE, "large_cases.json"), "w") as f:
        json.dump(large, f, indent=1)

    # Jacobian-transpose products from impulse responses
    jac = {}
    jcases = [("o0_mirror_64", (64, 64), [3.0], [0], "mirror"),
              ("o1_mirror_64", (64, 64), [3.0], [1, 0], "mirror"),
              ("o0_periodic_64", (64, 64), [3.0], [0], "periodic"),
              ("o1_periodic_64", (64, 64), [3.0], [0, 1], "periodic"),
              ("o0_mirror_3d12", (12, 12, 12), [1.5], [0], "mirror")]
    for k, (name, shape, sig, ords, bc) in enumerate(jcases):
        w = np.random.default_rng(500 + k).standard_normal(shape)
        n = int(np.prod(shape))
        jtw = np.zeros(n)
        e = np.zeros(shape)
        ef = e.reshape(-1)
        for j in range(n):
            ef[j] = 1.0
            resp = oc.gauss_fir(e, sig, ords, [bc], 3.0, use_ref=True)
            jtw[j] = float((resp * w).sum())
            ef[j] = 0.0
        jac[name + "__w"] = w
        jac[name + "__jtw"] = jtw.reshape(shape)
        jac[name + "__meta"] = np.array(json.dumps({"shape": shape, "sigmas": sig, "orders": ords, "bc": bc}))
        print("jacobian", name, flush=True)
    np.savez_compressed(os.path.join(HERE, "jacobian.npz"), **jac)


if __name__ == "__main__":
    main()
