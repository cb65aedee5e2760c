"""CPU checks of the boundary: the HIP library loads, exports every symbol
that include/qmctoolscl_amd.h declares, and the Python surface mirrors the
reference's argument lists and error behaviour without touching a GPU."""
import ctypes
import os
import subprocess
import sys
import re

import numpy as np
import pytest

import qmctoolscl_amd as q
from qmctoolscl_amd import _lib
from qmctoolscl_amd._api import SIGNATURES

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "qmctoolscl_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"QMC_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


def test_header_declares_every_reference_kernel():
    syms = set(declared_symbols())
    for name in SIGNATURES:
        assert name in syms and name + "_dev" in syms, name
    assert len(SIGNATURES) == 23


def test_library_loads_and_exports_all_declared_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (python -m qmctoolscl_amd.build)")
    L = _lib.lib()  # prototypes declare cleanly
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert missing == []
    assert b"gfx950" in L.qmc_version()


def test_library_shares_torchs_hip_runtime():
    """The package loads the library after torch, whose ROCm wheel bundles a
    libamdhip64 of the same soname: the process then maps one HIP runtime, so
    torch's streams are valid handles in the library.  Checked in a fresh
    interpreter that touches the library before importing torch itself."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (python -m qmctoolscl_amd.build)")
    pytest.importorskip("torch")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from qmctoolscl_amd import _lib\n"
            "_lib.lib()\n"
            "import torch\n"
            "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(maps), sorted(maps))\n") % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split()[0] == "1", out.stdout


def test_signatures_match_reference_prototypes_order():
    # argument names in reference order (qmctoolscl/c_funcs/<name>.c) spot checks
    assert [s[0] for s in SIGNATURES["gdn_digital_permutation"]] == [
        "r", "n", "d", "r_x", "r_b", "tmax", "tmax_new", "bmax", "perms", "xdig", "xdig_new"]
    assert [s[0] for s in SIGNATURES["dnb2_interlace"]] == [
        "r", "d_alpha", "mmax", "d", "tmax", "tmax_alpha", "alpha", "C", "C_alpha"]
    assert [s[0] for s in SIGNATURES["fft_bro_1d_radix2"]] == [
        "d1", "d2", "n_half", "twiddler", "twiddlei", "xr", "xi"]


def test_python_surface_errors_without_gpu():
    g = np.ones((1, 2), dtype=np.uint64)
    x = np.empty((1, 3, 2))
    # natural-order precondition of qmctoolscl/util.py:84-86
    with pytest.raises(Exception, match="powers of 2"):
        q.lat_gen_natural(np.uint64(1), np.uint64(3), np.uint64(2), np.uint64(0), g, x)
    # unknown backend, as util.py:54
    with pytest.raises(AssertionError):
        q.lat_gen_linear(np.uint64(1), np.uint64(3), np.uint64(2), g, x, backend="cuda")
    # wrong argument count
    with pytest.raises(TypeError):
        q.lat_gen_linear(np.uint64(1), np.uint64(3), g, x)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_dtype_checks_like_ndpointer():
    g = np.ones((1, 2), dtype=np.int64)  # wrong dtype
    x = np.empty((1, 3, 2))
    with pytest.raises(ctypes.ArgumentError):
        q.lat_gen_linear(np.uint64(1), np.uint64(3), np.uint64(2), g, x)
    g = np.ones((1, 4), dtype=np.uint64)[:, ::2]  # not C-contiguous
    with pytest.raises(ctypes.ArgumentError):
        q.lat_gen_linear(np.uint64(1), np.uint64(3), np.uint64(2), g, x)


def test_entry_points_reject_bad_arguments_before_any_gpu_call():
    """Argument checks come first: these calls return their error codes (and
    a message) on a machine with no GPU, as on one with."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (python -m qmctoolscl_amd.build)")
    L = _lib.lib()
    u64, vp, ci = ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int
    fused = L.qmc_gdn_gen_scramble_float_dev
    fused.restype = ci
    fused.argtypes = [ci] + [u64] * 9 + [vp] * 5
    ARG, RANGE = 1, 2
    #            mode  r  n  d  r_b mmax tmax tnew n_start bmax
    cases = [((3, 1, 8, 2, 1, 4, 4, 4, 0, 5), ARG),    # no such mode
             ((1, 1, 8, 2, 1, 4, 5, 4, 0, 5), ARG),    # tmax > tmax_new
             ((1, 1, 8, 2, 1, 4, 4, 65, 0, 5), ARG),   # tmax_new above 64
             ((0, 1, 8, 2, 1, 4, 4, 6, 0, 5), ARG),    # mode 0 needs tmax_new == tmax
             ((1, 1, 8, 2, 1, 65, 4, 4, 0, 5), ARG),   # mmax above 64
             ((1, 1, 8, 2, 0, 4, 4, 4, 0, 5), ARG),    # r_b = 0
             ((1, 1, 8, 2, 1, 4, 4, 4, 0, 0), ARG),    # bmax = 0 with permutations
             ((1, 0, 8, 2, 1, 4, 4, 4, 0, 5), 0)]      # nothing to do
    for args, want in cases:
        assert fused(*args, None, None, None, None, None) == want, args
        if want:
            assert L.qmc_last_error().decode()
    halton = L.qmc_halton_qrng_dev
    halton.restype = ci
    halton.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp]
    assert halton(-1, 2, 0, 1, None, None, None, None) == ARG
    assert halton(10, 2, 2**31 - 5, 1, None, None, None, None) == RANGE
    assert halton(0, 2, 0, 1, None, None, None, None) == 0
    fw = L.qmc_fwht_mfma_dev
    fw.restype = ci
    fw.argtypes = [u64, u64, u64, vp, vp, vp]
    assert fw(1, 1, 384, None, None, None) == ARG     # 768 is not a power of two
    assert fw(1, 1, 64, None, None, None) == ARG      # 2^7 is below the minimum
    assert fw(1, 1, 1 << 24, None, None, None) == ARG  # 2^25 is above the maximum
    assert fw(1, 1, 512, 8, None, None) == ARG         # misaligned pointer


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_undersized_arrays_are_rejected_before_any_call():
    """Every array argument must hold the elements the call touches (the
    sizes of the reference's prototype comments); a smaller one raises
    ctypes.ArgumentError before the library is asked to do anything."""
    from qmctoolscl_amd._api import SIZES, required_elements
    U = np.uint64
    g = np.ones((2, 3), dtype=np.uint64)
    with pytest.raises(ctypes.ArgumentError, match="x holds 17"):
        q.lat_gen_linear(U(2), U(3), U(3), g, np.empty(17))        # needs 2*3*3 = 18
    with pytest.raises(ctypes.ArgumentError, match="g holds 5"):
        q.lat_gen_linear(U(2), U(3), U(3), g.ravel()[:5].copy(), np.empty(18))
    with pytest.raises(ctypes.ArgumentError, match="perms"):
        q.gdn_digital_permutation(U(1), U(4), U(2), U(1), U(1), U(3), U(3), U(5), np.zeros(29, np.uint64),
                                  np.zeros(24, np.uint64), np.zeros(24, np.uint64))  # perms needs 1*2*3*5 = 30
    with pytest.raises(ctypes.ArgumentError, match="lshifts"):
        q.dnb2_digital_shift(U(4), U(2), U(2), U(3), np.zeros(2, np.uint64), np.zeros(12, np.uint64),
                             np.zeros(8, np.uint64), np.zeros(16, np.uint64))  # lshifts[l % r_x]: r_x = 3
    with pytest.raises(ctypes.ArgumentError, match="xi"):
        q.fft_bro_1d_radix2(U(2), U(1), U(4), np.empty(8), np.empty(8), np.empty(16), np.empty(15))
    # a spot check of the table against the reference's index arithmetic
    assert required_elements("gdn_linear_matrix_scramble", dict(r=3, d=2, mmax=4, r_C=1, r_b=1, tmax=5,
                                                                tmax_new=7)) == {
        "bases": 2, "S": 3 * 2 * 7 * 5, "C": 1 * 2 * 4 * 5, "C_lms": 3 * 2 * 4 * 7}
    assert set(SIZES) == set(SIGNATURES)
