"""Shared helpers for the parity tests (comparison metrics, device staging)."""
import numpy as np

PHASE_TOL = 1.0e-3   # rad, wrapped phase vs the FP64 oracle
TCORR_TOL = 1.0e-4   # temporal coherence vs the FP64 oracle
GAP_MIN = 1.0e-3     # relative eigen-gap below which a pixel's phase is ill-posed
# Smallest eigenvector component, relative to the RMS one, whose phase an FP32
# solver fixes to 1e-3 rad: its absolute eigenvector error is ~1e-7, so a
# component of size |v_t| carries a phase error of ~1e-7 / |v_t| (FP64 fixes
# such phases; FP32 in general does not).
VMIN_MIN = 1.0e-2


def wrapped_err(a, b):
    """|arg(a conj(b))| elementwise (complex arrays)."""
    return np.abs(np.angle(a * np.conj(b)))


def to_dev(x, torch, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def wts_dev(wts, torch):
    return torch.from_numpy(np.ascontiguousarray(wts).view(np.int32)).to("cuda")


def compare_link(gpu, ref, gap=None, phase_tol=PHASE_TOL, tcorr_tol=TCORR_TOL, gap_min=GAP_MIN, vmin=None,
                 vmin_min=VMIN_MIN, stack=None, ref_index=0):
    """Return a dict of mismatch statistics between GPU and oracle outputs.
    Pixels count as well posed when processed, with eigen-gap >= gap_min and
    (when vmin is given) eigenvector components no smaller than vmin_min of
    the RMS component.  With the stack given, comp_err is the compressed-SLC
    error over the mean |z| of the dates it sums (evd.cpp:755-762): at most
    the largest phase error, plus rounding."""
    ph_g, tc_g, co_g = gpu
    ph_r, tc_r, co_r = ref["phase"], ref["tcorr"], ref["comp"]
    ok_r = tc_r > 0            # processed, not failed
    well = ok_r.copy()
    if gap is not None:
        well &= gap >= gap_min
    if vmin is not None:
        well &= vmin >= vmin_min
    err = wrapped_err(ph_g, ph_r).max(axis=0)
    out = {
        "n_ok": int(ok_r.sum()), "n_well": int(well.sum()),
        "phase_max": float(err[well].max()) if well.any() else 0.0,
        "tcorr_max": float(np.abs(tc_g - tc_r)[well].max()) if well.any() else 0.0,
        "codes_equal": bool(np.array_equal(np.where(tc_r <= 0, tc_r, 1), np.where(tc_g <= 0, tc_g, 1))),
        "untouched_equal": bool(np.array_equal(tc_r == 0, tc_g == 0)),
    }
    # compressed SLC relative error on well-posed pixels
    den = np.maximum(np.abs(co_r), 1e-30)
    out["comp_rel_max"] = float((np.abs(co_g - co_r) / den)[well].max()) if well.any() else 0.0
    if stack is not None:
        amp = np.abs(stack[ref_index:]).astype(np.float64).mean(axis=0)
        e = np.abs(co_g.astype(np.complex128) - co_r) / np.maximum(amp, 1e-30)
        out["comp_err"] = float(e[well].max()) if well.any() else 0.0
    return out
