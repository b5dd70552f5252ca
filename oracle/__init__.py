"""CPU oracle for fringe's phase-linking hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The GPU
library (``fringe_amd``) never imports, links or calls anything here.

``liboracle.so`` is built from ``fringe_oracle.cpp`` in this folder (plain C++,
OpenMP, FP64 LAPACK resolved at run time).  ``_ref/libfringe_ref.so`` is the
reference's own header-only kernels (KS2sample.hpp, AD2unique.hpp,
ulongmask.hpp, EigenLapack.hpp) compiled by ``Makefile`` where the reference
tree is present; it pins the restatement.

Parity status of each oracle function (what pins it, see tests/):
  kolmogorov_prob, ks2_pvalue   -> reference KS2sample.hpp via _ref (exact),
                                   exact rational D from brute-force ECDFs
  nmap                          -> brute-force numpy window sweep + _ref KS
  eigen_index, pd_inverse       -> reference EigenLapack.hpp via _ref, the
                                   3x3 known-answer case of tests/eigen/test_eig.cpp
  calamp                        -> closed forms (constant amplitude, empty mask,
                                   NaN and zero pixels excluded)
  despeck                       -> closed forms (constant fields, single-bit
                                   masks, coherence of a band with itself)
  link (Gram/coherence)         -> numpy FP64 brute force on tiny windows
  link (EVD/MLE/EMI)            -> closed forms (Perron-Frobenius, EMI
                                   lambda=1 / v=exp(i theta)), numpy eigh
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

MOD_EVD, MOD_PHASE_LINK = 0, 1
METHODS = {"MLE": 0, "EVD": 1, "STBAS": 2}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so (and _ref when the reference tree exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_lapack_load.argtypes = [ctypes.c_char_p]
        L.oracle_mask_words.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_kolmogorov_prob.argtypes = [ctypes.c_double]
        L.oracle_kolmogorov_prob.restype = ctypes.c_double
        L.oracle_ks2_statistic.argtypes = [_f32p, _f32p, ctypes.c_int]
        L.oracle_ks2_statistic.restype = ctypes.c_double
        L.oracle_ks2_pvalue.argtypes = [_f32p, _f32p, ctypes.c_int]
        L.oracle_ks2_pvalue.restype = ctypes.c_double
        L.oracle_nmap.argtypes = [_f32p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, _u32p, _i32p, ctypes.c_int]
        L.oracle_calamp.argtypes = [_f32p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_despeck.argtypes = [_f32p, ctypes.c_void_p, _u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _f32p]
        L.oracle_ad2_statistic.argtypes = [_f32p, _f32p, ctypes.c_int]
        L.oracle_ad2_statistic.restype = ctypes.c_double
        L.oracle_ad2_pvalue.argtypes = [_f32p, _f32p, ctypes.c_int]
        L.oracle_ad2_pvalue.restype = ctypes.c_double
        L.oracle_ad2_pvalue_of.argtypes = [ctypes.c_double]
        L.oracle_ad2_pvalue_of.restype = ctypes.c_double
        L.oracle_ad2_sigma.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_ad2_sigma.restype = ctypes.c_double
        L.oracle_link.argtypes = [_f32p, _u32p] + [ctypes.c_int] * 12 + [
            _f32p, _f32p, _f32p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_eigen_index.argtypes = [_f64p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
        L.oracle_pd_inverse.argtypes = [_f64p, ctypes.c_int]
        _LIB = L
    return _LIB


def ref():
    """The reference's header kernels (None where _ref was not built)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libfringe_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_ks2_test.argtypes = [_f32p, _f32p, ctypes.c_int]
        R.ref_ks2_test.restype = ctypes.c_double
        R.ref_kolmogorov_prob.argtypes = [ctypes.c_double]
        R.ref_kolmogorov_prob.restype = ctypes.c_double
        R.ref_ad2_test.argtypes = [_f32p, _f32p, ctypes.c_int]
        R.ref_ad2_test.restype = ctypes.c_double
        R.ref_setbit.argtypes = [_u32p] + [ctypes.c_int] * 4
        R.ref_getbit.argtypes = [_u32p] + [ctypes.c_int] * 4
        R.ref_eigen_index.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), _f64p]
        R.ref_pd_inverse.argtypes = [_f64p, ctypes.c_int]
        _REF = R
    return _REF


def mask_words(Nx, Ny):
    return lib().oracle_mask_words(Nx, Ny)


def kolmogorov_prob(z):
    return lib().oracle_kolmogorov_prob(float(z))


def ks2_statistic(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().oracle_ks2_statistic(a, b, a.size)


def ad2_statistic(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().oracle_ad2_statistic(a, b, a.size)


def ad2_pvalue(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().oracle_ad2_pvalue(a, b, a.size)


def ad2_pvalue_of(t):
    return lib().oracle_ad2_pvalue_of(float(t))


def ks2_pvalue(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().oracle_ks2_pvalue(a, b, a.size)


def _slc_f32(slc):
    slc = np.ascontiguousarray(slc, np.complex64)
    return slc.view(np.float32).reshape(-1)


def nmap(slc, Nx, Ny, prob=0.05, mask=None, amp_const=None, threads=0, method="KS2"):
    """Race-free SHP map (method "KS2" or "AD2").  slc: complex64 [nbands, rows, cols].

    Returns (wts uint32 [rows, cols, nulong], count int32 [rows, cols])."""
    nb, rows, cols = slc.shape
    nul = mask_words(Nx, Ny)
    wts = np.zeros((rows, cols, nul), np.uint32)
    cnt = np.zeros((rows, cols), np.int32)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    ac = None if amp_const is None else np.ascontiguousarray(amp_const, np.float64)
    lib().oracle_nmap(_slc_f32(slc), None if m is None else m.ctypes.data, rows, cols, nb,
                      None if ac is None else ac.ctypes.data, Nx, Ny, float(prob),
                      {"KS2": 0, "AD2": 1}[method], wts.reshape(-1), cnt.reshape(-1), int(threads))
    return wts, cnt


def calamp(slc, mask=None):
    """Per-band mean |z| over valid pixels (calamp.cpp:180-262)."""
    nb, rows, cols = slc.shape
    out = np.zeros(nb, np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    lib().oracle_calamp(_slc_f32(slc), None if m is None else m.ctypes.data, rows, cols, nb, out)
    return out


def despeck(b1, wts, Nx, Ny, b2=None, coherence=False):
    """SHP-masked mean (despeck.cpp:300-433): amplitude of b1, or the
    interferogram b1 conj(b2), or its coherence.  Returns complex64
    [rows, cols] (the amplitude is the real part)."""
    rows, cols = b1.shape
    mode = 0 if b2 is None else (2 if coherence else 1)
    out = np.zeros((rows, cols), np.complex64)
    a = np.ascontiguousarray(b1, np.complex64)
    b = None if b2 is None else np.ascontiguousarray(b2, np.complex64)
    lib().oracle_despeck(a.view(np.float32).reshape(-1), None if b is None else b.ctypes.data,
                         np.ascontiguousarray(wts, np.uint32).reshape(-1), rows, cols, Nx, Ny, mode,
                         out.view(np.float32).reshape(-1))
    return out


def link(slc, wts, Nx, Ny, module="phase_link", method="MLE", min_neighbors=2,
         mini_stack_count=1, bandwidth=-1, rows_range=None, want_gap=False, threads=0):
    """Per-pixel phase linking with reference semantics.

    Returns dict(phase=[nbands,rows,cols] c64, tcorr=[rows,cols] f32,
    comp=[rows,cols] c64, status=int[, gap=[rows,cols] f64, vmin=[rows,cols] f64]).
    gap: eigen-gap of the selected eigenvalue over the spectral spread;
    vmin: sqrt(n) times the smallest modulus of the selected unit eigenvector."""
    nb, rows, cols = slc.shape
    r0, r1 = (0, rows) if rows_range is None else rows_range
    phase = np.zeros((nb, rows, cols), np.complex64)
    tcorr = np.zeros((rows, cols), np.float32)
    comp = np.zeros((rows, cols), np.complex64)
    gap = np.full((rows, cols), -1.0) if want_gap else None
    vmin = np.full((rows, cols), -1.0) if want_gap else None
    mod = MOD_EVD if module == "evd" else MOD_PHASE_LINK
    st = lib().oracle_link(_slc_f32(slc), np.ascontiguousarray(wts, np.uint32).reshape(-1),
                           rows, cols, nb, Nx, Ny, mod, METHODS[method], int(min_neighbors),
                           int(mini_stack_count), int(bandwidth), r0, r1,
                           phase.view(np.float32).reshape(-1), tcorr.reshape(-1),
                           comp.view(np.float32).reshape(-1),
                           None if gap is None else gap.ctypes.data,
                           None if vmin is None else vmin.ctypes.data, int(threads))
    out = dict(phase=phase, tcorr=tcorr, comp=comp, status=st)
    if want_gap:
        out["gap"] = gap
        out["vmin"] = vmin
    return out


def eigen_index(A, index, want_vec=True):
    """zheevr RANGE='I' single eigenpair of Hermitian A (uses the upper triangle)."""
    n = A.shape[0]
    a = np.asfortranarray(A, np.complex128).ravel(order="F").view(np.float64).copy()
    val = ctypes.c_double()
    vec = np.zeros(2 * n, np.float64) if want_vec else None
    info = lib().oracle_eigen_index(a, n, int(index), ctypes.byref(val),
                                    None if vec is None else vec.ctypes.data)
    return info, val.value, (None if vec is None else vec.view(np.complex128))


def pd_inverse(A):
    n = A.shape[0]
    a = np.asfortranarray(A, np.complex128).ravel(order="F").view(np.float64).copy()
    info = lib().oracle_pd_inverse(a, n)
    return info, a.view(np.complex128).reshape((n, n), order="F")
