"""GPU parity on the shapes of the headline configurations that the smaller
parity tests do not reach: nmap at the 15 x 11 and 21 x 11 windows (and the
reference sequential default 59 x 19) with invalid pixels, a byte mask and
amplitude constants; the N = 64, 21 x 11 Gram and EVD on a full 4096-column
band; the limits of the solver above 64 dates."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available
from helpers import PHASE_TOL, TCORR_TOL, compare_link, to_dev, wts_dev

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

import oracle  # noqa: E402
from fringe_amd import _lib, synthetic, tile  # noqa: E402


def _with_invalid_pixels(slc, rng):
    """A zero sample, a NaN sample and a whole zero pixel (nmap.cpp:345-382:
    any band with |z| == 0 or NaN invalidates the pixel)."""
    n, rows, cols = slc.shape
    out = slc.copy()
    for _ in range(3):
        out[rng.integers(n), rng.integers(rows), rng.integers(cols)] = 0
    for _ in range(3):
        out[rng.integers(n), rng.integers(rows), rng.integers(cols)] = complex(np.nan, 0.5)
    out[:, rows // 2, cols // 3] = 0
    return out


@pytest.mark.parametrize("method", ["KS2", "AD2"])
@pytest.mark.parametrize("n,Nx,Ny", [(30, 7, 5), (30, 10, 5), (10, 29, 9)])
def test_nmap_bit_exact_large_windows_invalid_pixels_mask(n, Nx, Ny, method):
    rows, cols = (24, 40) if Nx < 20 else (26, 70)
    print(f"case nmap n={n} window {2 * Nx + 1}x{2 * Ny + 1} {method} with mask", flush=True)
    rng = np.random.default_rng(700 + n + Nx)
    slc = _with_invalid_pixels(synthetic.stack(n, rows, cols, seed=700 + Nx, patches=6), rng)
    mask = (rng.random((rows, cols)) > 0.05).astype(np.uint8)     # maskDS: 0 = unusable
    amp = np.ones(n)
    amp[1], amp[2], amp[4] = 2.0, -1.0, 0.0                        # <= 0 reads as missing
    # scale band 1 by its constant: with alpha_b the maps are unchanged
    slc[1] *= 2.0
    wr, cr = oracle.nmap(slc, Nx, Ny, 0.05, mask=mask, amp_const=amp, method=method)
    wg, cg = tile.nmap(to_dev(slc, torch), Nx, Ny, 0.05, mask=torch.from_numpy(mask), amp_const=amp,
                       method=method)
    torch.cuda.synchronize()
    wg = wg.cpu().numpy().view(np.uint32)
    cg = cg.cpu().numpy()
    assert np.array_equal(wg, wr)
    assert np.array_equal(cg, cr)
    # the masked and the invalid pixels get no bits and count 0
    assert not wg[mask == 0].any() and not cg[mask == 0].any()
    assert cg[rows // 2, cols // 3] == 0
    assert (cr > 1).mean() > 0.5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_link_sequential_default_window_59x19(precision):
    # sequential.py's default half-window (29, 9): 36 mask words
    n, rows, cols, Nx, Ny = 10, 26, 70, 29, 9
    print(f"case link 59x19 {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=759, patches=8)
    wts, _ = oracle.nmap(slc, Nx, Ny, 0.05)
    r = oracle.link(slc, wts, Nx, Ny, module="evd", method="MLE", min_neighbors=5, want_gap=True)
    ph, tc, co = tile.link(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, module="evd", method="MLE",
                           min_neighbors=5, precision=precision)
    torch.cuda.synchronize()
    st = compare_link((ph.cpu().numpy(), tc.cpu().numpy(), co.cpu().numpy()), r, r["gap"])
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["n_well"] > 0.9 * rows * cols
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL


@pytest.mark.parametrize("precision", ["mixed", "mixed_bf16"])
def test_cfg3_band_gram_and_evd(precision):
    """N = 64, 21 x 11 all-ones window, evd EVD: two output rows of a
    4096-column band (with their Ny halo rows).  The hi/lo Gram's coherence
    matrices against the FP64 Gram (itself against a numpy FP64 brute force
    on a few pixels), and the top-eigenvector phases against the oracle."""
    n, Nx, Ny, cols, out_rows = 64, 10, 5, 4096, 2
    rows = out_rows + 2 * Ny
    rb, re = Ny, Ny + out_rows
    print(f"case cfg3 band n={n} {rows}x{cols} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=103)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    d_slc, d_wts = to_dev(slc, torch), wts_dev(wts, torch)
    g64, _ = tile.coherence(d_slc, d_wts, Nx, Ny, precision="fp64", module="evd", row_begin=rb, row_end=re)
    glo, _ = tile.coherence(d_slc, d_wts, Nx, Ny, precision=precision, module="evd", row_begin=rb, row_end=re)
    torch.cuda.synchronize()
    g64, glo = g64.cpu().numpy(), glo.cpu().numpy()
    for k in (0, 17, cols - 1, cols + 2049):
        r, c = rb + k // cols, k % cols
        Z = slc[:, max(r - Ny, 0):r + Ny + 1, max(c - Nx, 0):c + Nx + 1].reshape(n, -1).astype(np.complex128)
        G = Z @ Z.conj().T
        d = np.sqrt(np.real(np.diag(G)))
        assert np.abs(g64[k] - G / np.outer(d, d)).max() < 1e-6
    fro = np.sqrt((np.abs(glo - g64) ** 2).sum(axis=(1, 2)))
    print(f"  Frobenius |Gamma - Gamma_FP64|: p50 {np.median(fro):.2e} max {fro.max():.2e}")
    assert fro.max() < (1e-5 if precision == "mixed" else 2e-4)
    ref = oracle.link(slc, wts, Nx, Ny, module="evd", method="EVD", rows_range=(rb, re), want_gap=True)
    ph, tc, co = tile.link(d_slc, d_wts, Nx, Ny, module="evd", method="EVD", precision=precision,
                           row_begin=rb, row_end=re)
    torch.cuda.synchronize()
    g = (ph.cpu().numpy()[:, rb:re], tc.cpu().numpy()[rb:re], co.cpu().numpy()[rb:re])
    rr = {k: (v[:, rb:re] if k == "phase" else v[rb:re]) for k, v in ref.items() if k != "status"}
    st = compare_link(g, rr, rr["gap"])
    print(" ", st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["n_well"] >= 0.99 * out_rows * cols
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL


def test_solver_limits_above_64_dates():
    """Above 64 dates: a date count beyond the solver's limit (256) is refused
    before any launch (code 202, no workspace); FP64 cases whose work matrix
    does not fit a workgroup's LDS (100 and 160 dates) run with it in global
    memory, and 128 dates in mixed precision run; all match the oracle."""
    L = _lib.lib()
    p = tile.link_params(16, 20, 257, 7, 5, precision="fp64")
    assert L.fringe_link_workspace_bytes(ctypes.byref(p)) == 0
    dummy = ctypes.c_void_p(256)
    rc = L.fringe_link_run(ctypes.byref(p), dummy, dummy, dummy, dummy, dummy, dummy, 1 << 20, None)
    assert rc == 202
    for n, rows, cols, Nx, Ny, precision in ((100, 14, 16, 7, 5, "fp64"), (160, 20, 22, 9, 6, "fp64"),
                                             (128, 14, 16, 7, 5, "mixed")):
        print(f"case {precision} n={n}", flush=True)
        slc = synthetic.stack(n, rows, cols, seed=1000 + n)
        wts = synthetic.full_mask(rows, cols, Nx, Ny)
        r = oracle.link(slc, wts, Nx, Ny, module="phase_link", method="MLE", want_gap=True)
        ph, tc, co = tile.link(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, module="phase_link", method="MLE",
                               precision=precision)
        torch.cuda.synchronize()
        st = compare_link((ph.cpu().numpy(), tc.cpu().numpy(), co.cpu().numpy()), r, r["gap"], stack=slc)
        print(st)
        assert st["untouched_equal"] and st["codes_equal"]
        assert st["n_well"] >= 16
        assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL
