"""CPU-side checks of the C ABI library (no GPU needed): it loads, exports every
entry point declared in include/dipamd.h, computes the host-side taps exactly as
the reference, and rejects bad arguments before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_ctypes as oc
from diplib_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TAPS = np.load(os.path.join(ROOT, "tests", "golden", "taps.npz"))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(dipamd_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) >= set(_lib.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("key", list(TAPS.keys()))
def test_host_taps_match_reference(key):
    m = re.match(r"s([0-9.]+)_o(\d)_d(\d)", key)
    s, o, d = float(m.group(1)), int(m.group(2)), int(m.group(3))
    buf = (ctypes.c_double * 4096)()
    n = _lib.lib().dipamd_make_half_gaussian(s, o, 3.0, d, buf, 4096)
    assert n == len(TAPS[key])
    assert np.array_equal(np.frombuffer(buf, dtype=np.float64, count=n), TAPS[key])


def test_radius_matches_oracle():
    L = _lib.lib()
    for s in [0.5, 1, 2, 3, 4.5, 8, 16, 33.3]:
        for o in range(4):
            for d in (0, 1):
                assert L.dipamd_gauss_radius(s, o, 3.0, d) == oc.gauss_radius(s, o, 3.0, bool(d))


def _img(ptr, dt, sizes):
    im = _lib.DipamdImage()
    im.origin = ptr
    im.dtype = dt
    im.ndim = len(sizes)
    st = 1
    for i, s in enumerate(sizes):
        im.sizes[i] = s
        im.strides[i] = st
        st *= s
    return im


def test_argument_errors_without_device():
    """Validation happens before any HIP call, so these run on a CPU-only host."""
    L = _lib.lib()
    sig = (ctypes.c_double * 1)(1.0)
    o = (ctypes.c_int32 * 1)(4)
    good = _img(ctypes.c_void_p(0x1000), _lib.DT_SFLOAT, [8, 8])
    unforged = _img(None, _lib.DT_SFLOAT, [8, 8])
    assert L.dipamd_gauss_fir(ctypes.byref(unforged), ctypes.byref(good), sig, 1, None, 0, None, 0, 3.0, 0,
                              None) == _lib.E_NOT_FORGED
    assert b"not forged" in L.dipamd_last_error()
    assert L.dipamd_gauss_fir(ctypes.byref(good), ctypes.byref(good), sig, 1, o, 1, None, 0, 3.0, 0,
                              None) == _lib.E_ORDER
    sig3 = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    assert L.dipamd_gauss_fir(ctypes.byref(good), ctypes.byref(good), sig3, 3, None, 0, None, 0, 3.0, 0,
                              None) == _lib.E_ARRAY_LENGTH
    bad_bc = (ctypes.c_int32 * 1)(99)
    assert L.dipamd_gauss_fir(ctypes.byref(good), ctypes.byref(good), sig, 1, None, 0, bad_bc, 1, 3.0, 0,
                              None) == _lib.E_PARAMETER
    other = _img(ctypes.c_void_p(0x1000), _lib.DT_SFLOAT, [8, 9])
    f = _lib.DipamdFilter()
    taps = (ctypes.c_double * 3)(1, 2, 1)
    f.taps = ctypes.cast(taps, ctypes.POINTER(ctypes.c_double))
    f.ntaps = 3
    f.origin = -1
    f.symmetry = 0
    assert L.dipamd_separable_convolution(ctypes.byref(good), ctypes.byref(other), ctypes.byref(f), 1, None, 0,
                                          None, 0, None) == _lib.E_SIZES
    f.origin = 7
    assert L.dipamd_separable_convolution(ctypes.byref(good), ctypes.byref(good), ctypes.byref(f), 1, None, 0,
                                          None, 0, None) == _lib.E_PARAMETER
    assert b"Origin" in L.dipamd_last_error()
    # complex: a complex filter needs a complex output image; an odd-length complex
    # filter is refused when the Python layer builds it
    f.origin = -1
    f.is_complex = 1
    f.ntaps = 1
    assert L.dipamd_separable_convolution(ctypes.byref(good), ctypes.byref(good), ctypes.byref(f), 1, None, 0,
                                          None, 0, None) == _lib.E_UNSUPPORTED
    assert b"complex output" in L.dipamd_last_error()
    from diplib_amd import OneDimensionalFilter
    with pytest.raises(_lib.DipError):
        OneDimensionalFilter([1.0, 2.0, 3.0], isComplex=True)
    cf = OneDimensionalFilter([1 + 2j, -0.5j])
    assert cf.isComplex and cf.filter == [1.0, 2.0, -0.0, -0.5]


def test_line_buffer_argument_errors_without_device():
    """The plug-in framework's line-buffer entry points validate before any HIP call."""
    L = _lib.lib()
    L.dipamd_separable_lines_count.restype = ctypes.c_int64
    good = _img(ctypes.c_void_p(0x1000), _lib.DT_SFLOAT, [8, 6, 5])
    assert L.dipamd_separable_lines_count(ctypes.byref(good), 1) == 40
    assert L.dipamd_separable_lines_count(ctypes.byref(good), 3) == _lib.E_PARAMETER
    buf = ctypes.c_void_p(0x2000)
    assert L.dipamd_separable_lines_gather(ctypes.byref(good), 0, 2, 0, _lib.DT_UINT8, buf, None) == _lib.E_UNSUPPORTED
    assert L.dipamd_separable_lines_gather(ctypes.byref(good), 0, -1, 0, _lib.DT_SFLOAT, buf, None) == _lib.E_PARAMETER
    assert L.dipamd_separable_lines_gather(ctypes.byref(good), 0, 2, 99, _lib.DT_SFLOAT, buf, None) == _lib.E_PARAMETER
    assert L.dipamd_separable_lines_scatter(None, _lib.DT_SFLOAT, ctypes.byref(good), 0, None) == _lib.E_PARAMETER
    assert L.dipamd_separable_lines_scatter(buf, _lib.DT_SFLOAT, ctypes.byref(good), 5, None) == _lib.E_PARAMETER


def test_framework_header_compiles(tmp_path):
    """include/diplib_amd_framework.hpp compiles for gfx950 with a user line filter (no GPU needed)."""
    import subprocess
    src = tmp_path / "t.hip"
    src.write_text('#include "diplib_amd_framework.hpp"\n'
                   'struct Copy { __device__ void Filter(dip::Framework::SeparableLineFilterParameters const &p) '
                   'const { float const *i = static_cast<float const *>(p.inBuffer.buffer); '
                   'float *o = static_cast<float *>(p.outBuffer.buffer); '
                   'for (dip::uint k = 0; k < p.inBuffer.length; ++k) o[k * p.outBuffer.stride] = '
                   'i[k * p.inBuffer.stride]; } };\n'
                   'void f(dip::Image const &a, dip::Image &b) { dip::Framework::Separable(a, b, dip::DT_SFLOAT, '
                   'dip::DT_SFLOAT, {}, {1}, {"mirror"}, Copy{}); }\n')
    p = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "--offload-arch=gfx950", "-c", "-I",
                        os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "t.o")],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_best_method_resolution():
    from diplib_amd import best_gauss_method
    assert best_gauss_method([2.0], [0]) == "fir"
    assert best_gauss_method([0.5], [0]) == "ft"
    assert best_gauss_method([12.0], [0]) == "iir"
    assert best_gauss_method([2.0], [4]) == "ft"
    assert best_gauss_method([0.0, 2.0], [0]) == "fir"


def test_cpp_drop_in_header_compiles(tmp_path):
    """include/diplib_amd.hpp compiles as plain C++14 with g++ (no GPU needed)."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text('#include "diplib_amd.hpp"\n'
                   'int main() { dip::Image a; dip::Image b = dip::Gauss(a, {2.0}); (void)b; return 0; }\n')
    p = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I",
                        os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
