"""GPU parity: libfringe_amd.so (through its C ABI) against the CPU oracle.

Sizes are small enough for the FP64 LAPACK oracle to finish in seconds.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import compare_link, wts_dev, to_dev, PHASE_TOL, TCORR_TOL

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

import oracle  # noqa: E402
from fringe_amd import synthetic, tile  # noqa: E402


def _run_both(slc, wts, Nx, Ny, module, method, precision, min_nb=2, ref=0, bw=-1, want_gap=True,
              want_resolved=False):
    r = oracle.link(slc, wts, Nx, Ny, module=module, method=method, min_neighbors=min_nb,
                    mini_stack_count=ref + 1, bandwidth=bw, want_gap=want_gap)
    n, rows, cols = slc.shape
    runner = tile.LinkRunner(tile.link_params(rows, cols, n, Nx, Ny, module=module, method=method,
                                              min_neighbors=min_nb, ref_index=ref, bandwidth=bw,
                                              precision=precision))
    ph, tc, co = runner.run(to_dev(slc, torch), wts_dev(wts, torch))
    torch.cuda.synchronize()
    g = (ph.cpu().numpy(), tc.cpu().numpy(), co.cpu().numpy())
    if want_resolved:
        return g, r, runner.resolved_pixels()
    return g, r


class _GenericPlace:
    """Sets FRINGE_SCHED_GENERIC_A (the solver's layout above 32 dates: rows
    in registers, one row per lane up to 64 dates, or a work matrix in LDS or
    global memory above) for one key, restoring it."""

    def __init__(self, n, precision, place):
        from fringe_amd import _lib
        self.L, self.key = _lib.lib(), (n, 0 if precision == "fp64" else 1, 0)
        self.v = None if place is None else _lib.GENERIC_A_PLACES[place]

    def __enter__(self):
        from fringe_amd import _lib
        self.old = self.L.fringe_schedule_get(_lib.FRINGE_SCHED_GENERIC_A, *self.key)
        if self.v is not None:
            assert self.L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, *self.key, self.v) == 0

    def __exit__(self, *exc):
        from fringe_amd import _lib
        self.L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, *self.key, self.old)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("module,method", [("phase_link", "MLE"), ("evd", "EVD"), ("evd", "MLE")])
def test_link_full_mask_small(precision, module, method):
    n, rows, cols, Nx, Ny = 10, 20, 24, 3, 2
    print(f"case n={n} {rows}x{cols} win {Nx}x{Ny} {module}/{method} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=101)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    g, r = _run_both(slc, wts, Nx, Ny, module, method, precision)
    st = compare_link(g, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["n_well"] > 0.9 * rows * cols
    assert st["phase_max"] <= PHASE_TOL
    assert st["tcorr_max"] <= TCORR_TOL
    assert st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_link_n30_nmap_mask(precision):
    n, rows, cols, Nx, Ny = 30, 24, 40, 7, 5
    print(f"case n={n} nmap mask {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=102, patches=8)
    wts, cnt = oracle.nmap(slc, Nx, Ny, 0.05)
    g, r = _run_both(slc, wts, Nx, Ny, "phase_link", "MLE", precision, min_nb=5)
    st = compare_link(g, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["phase_max"] <= PHASE_TOL
    assert st["tcorr_max"] <= TCORR_TOL
    assert st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("method", ["KS2", "AD2"])
@pytest.mark.parametrize("n", [8, 10, 30, 64])
def test_nmap_bit_exact(n, method):
    rows, cols, Nx, Ny = 20, 28, 5, 3
    print(f"case nmap n={n} {method}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=200 + n, patches=6)
    if n == 10:   # quantised amplitudes exercise the tie branches of the walks
        slc = (np.round(np.abs(slc) * 4) / 4 * np.exp(1j * np.angle(slc))).astype(np.complex64)
    wr, cr = oracle.nmap(slc, Nx, Ny, 0.05, method=method)
    wg, cg = tile.nmap(to_dev(slc, torch), Nx, Ny, 0.05, method=method)
    torch.cuda.synchronize()
    wg = wg.cpu().numpy().view(np.uint32)
    cg = cg.cpu().numpy()
    assert np.array_equal(wg, wr)
    assert np.array_equal(cg, cr)


@pytest.mark.parametrize("method", ["KS2", "AD2"])
@pytest.mark.parametrize("n", [30, 72])
def test_nmap_row_range(n, method):
    """A call over rows [row_begin, row_end) gives those rows of the
    full-tile map (their backward neighbours above row_begin included) and
    leaves every other row of the output untouched."""
    rows, cols, Nx, Ny = 24, 26, 4, 3
    slc = synthetic.stack(n, rows, cols, seed=300 + n, patches=6)
    wr, cr = oracle.nmap(slc, Nx, Ny, 0.05, method=method)
    d = to_dev(slc, torch)
    for rb, re in ((0, 5), (7, 16), (20, 24), (2, 3), (10, 10)):
        print(f"case nmap rows [{rb}, {re}) n={n} {method}", flush=True)
        wg, cg = tile.nmap(d, Nx, Ny, 0.05, method=method, row_begin=rb, row_end=re)
        torch.cuda.synchronize()
        wg = wg.cpu().numpy().view(np.uint32)
        cg = cg.cpu().numpy()
        assert np.array_equal(wg[rb:re], wr[rb:re])
        assert np.array_equal(cg[rb:re], cr[rb:re])
        outside = np.ones(rows, bool)
        outside[rb:re] = False
        assert not wg[outside].any() and not cg[outside].any()


@pytest.mark.parametrize("place", ["lanes", "registers"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("module,method", [("phase_link", "MLE"), ("evd", "EVD"), ("evd", "MLE")])
@pytest.mark.parametrize("n", [33, 48, 64])
def test_link_64row_solver(n, module, method, precision, place):
    # 32 < n <= 64: one pixel per wave, one row of every matrix per lane
    # ("lanes"), or one pixel per workgroup of 64 rows of two (FP32) or four
    # (FP64) threads ("registers"); an 11 x 9 window holds 99 samples in the
    # interior, fewer at the edges
    rows, cols, Nx, Ny = 14, 16, 5, 4
    print(f"case 64-row n={n} {module}/{method} {precision} {place}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=300 + n)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    # the mixed run's FP64 re-solve follows the FP64 layout
    with _GenericPlace(n, precision, place), _GenericPlace(n, "fp64", place):
        g, r = _run_both(slc, wts, Nx, Ny, module, method, precision)
    st = compare_link(g, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"]
    assert st["codes_equal"]   # evd MLE failure codes as the reference sets them
    # at least the 6 x 6 pixels with whole windows; evd MLE's thresholds code
    # most pixels at n = 64
    assert st["n_well"] >= (1 if (module, method) == ("evd", "MLE") else 36)
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL
    assert st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("place", ["lanes", "registers"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [40, 64])
def test_link_64row_solver_stbas_reference(n, precision, place):
    # banded coherence (STBAS) and a reference date other than the first, in
    # both solver layouts of 33..64 dates
    rows, cols, Nx, Ny = 14, 16, 5, 4
    print(f"case 64-row stbas n={n} {precision} {place}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=300 + n)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    with _GenericPlace(n, precision, place), _GenericPlace(n, "fp64", place):
        g, r = _run_both(slc, wts, Nx, Ny, "evd", "STBAS", precision, bw=10, ref=3)
    st = compare_link(g, r, r["gap"], stack=slc, ref_index=3)
    print(st)
    assert st["n_well"] >= 36 and st["codes_equal"]
    assert st["comp_err"] <= PHASE_TOL
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL
    assert np.all(g[0][3][r["tcorr"] > 0] == 1 + 0j)


@pytest.mark.parametrize("precision,place", [("fp64", None), ("mixed", "lds"), ("mixed", "registers")])
def test_link_generic_solver_stbas_reference(precision, place):
    n, rows, cols, Nx, Ny = 80, 14, 16, 7, 5
    print(f"case generic stbas n={n} {precision} {place}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=380)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    with _GenericPlace(n, precision, place):
        g, r = _run_both(slc, wts, Nx, Ny, "evd", "STBAS", precision, bw=12, ref=5)
    st = compare_link(g, r, r["gap"], stack=slc, ref_index=5)
    print(st)
    assert st["n_well"] >= 16 and st["codes_equal"]
    assert st["comp_err"] <= PHASE_TOL
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL
    assert np.all(g[0][5][r["tcorr"] > 0] == 1 + 0j)


def test_link_64row_solver_bf16():
    n, rows, cols, Nx, Ny = 40, 12, 14, 4, 3
    print("case 64-row n=40 mixed_bf16", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=103)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    g, r = _run_both(slc, wts, Nx, Ny, "phase_link", "MLE", "mixed_bf16")
    st = compare_link(g, r, r["gap"])
    print(st)
    assert st["untouched_equal"]
    # the bf16 hi/lo pair carries ~16 significant bits (fp16 with per-date
    # scaling ~22): its Gamma error is ~64x larger, so it gets 5x the bounds
    assert st["phase_max"] <= 5 * PHASE_TOL and st["tcorr_max"] <= 5 * TCORR_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_link_generic_solver_n72(precision):
    # above 64 dates the generic LDS-resident solver runs
    n, rows, cols, Nx, Ny = 72, 12, 14, 5, 4
    print(f"case generic n=72 {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=372)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    g, r = _run_both(slc, wts, Nx, Ny, "phase_link", "MLE", precision)
    st = compare_link(g, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("precision,place", [("fp64", None), ("mixed", "lds"), ("mixed", "registers")])
@pytest.mark.parametrize("module,method", [("phase_link", "MLE"), ("evd", "EVD"), ("evd", "MLE")])
@pytest.mark.parametrize("n", [72, 96, 112])
def test_link_generic_solver_modes(n, module, method, precision, place):
    # every estimator of the solvers above 64 dates: the LDS-resident work
    # matrix, and in the mixed precisions the four-wave solver with each row
    # in registers (48 columns per thread up to 96 dates, 64 above); a
    # 15 x 11 window holds 165 samples in the interior (6 x 6 pixels), fewer
    # at the edges, where the sample coherence of 96 dates can be rank
    # deficient
    rows, cols, Nx, Ny = 16, 20, 7, 5
    print(f"case generic n={n} {module}/{method} {precision} {place}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=400 + n)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    with _GenericPlace(n, precision, place):
        g, r = _run_both(slc, wts, Nx, Ny, module, method, precision)
    st = compare_link(g, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    # evd MLE's thresholds code every pixel at 112 dates (the oracle's codes,
    # asserted equal above)
    assert st["n_well"] >= ((1 if n <= 96 else 0) if (module, method) == ("evd", "MLE") else 36)
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [33, 49, 65, 81, 97, 113, 128])
def test_register_rows_column_counts(n, precision):
    """Date counts just past each column count of the register-row solvers
    (48 .. 128 columns in FP32; 12 .. 32 per thread in FP64) and at the last
    one: most of the last block of columns is padding, and (33, 65, 97)
    most rows of the 64-, 96- or 128-row workgroup are idle."""
    rows, cols, Nx, Ny = 16, 20, 7, 5
    print(f"case register rows n={n} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=800 + n)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    for module, method in (("phase_link", "MLE"), ("evd", "EVD")):
        with _GenericPlace(n, precision, "registers"), _GenericPlace(n, "fp64", "registers"):
            g, r = _run_both(slc, wts, Nx, Ny, module, method, precision)
        st = compare_link(g, r, r["gap"], stack=slc)
        print(module, method, st)
        assert st["untouched_equal"] and st["codes_equal"]
        assert st["n_well"] >= 4
        assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("n,Nx,Ny,module,method,place,place64", [
    (10, 29, 9, "evd", "MLE", None, None), (30, 7, 5, "phase_link", "MLE", None, None),
    (64, 5, 4, "evd", "MLE", None, None), (64, 5, 4, "evd", "EVD", None, None),
    (48, 5, 4, "phase_link", "MLE", None, None), (48, 5, 4, "phase_link", "MLE", None, "lanes"),
    (64, 5, 4, "evd", "EVD", "registers", "lanes"), (96, 7, 5, "phase_link", "MLE", "lds", None),
    (96, 7, 5, "phase_link", "MLE", "registers", None), (80, 6, 5, "evd", "MLE", "lds", None),
    (80, 6, 5, "evd", "MLE", "registers", None), (88, 6, 5, "evd", "EVD", "registers", None),
    (120, 7, 5, "phase_link", "MLE", "registers", None)])
def test_mixed_resolve_equals_fp64(n, Nx, Ny, module, method, place, place64):
    """Pixels whose FP32 eigenpair the mixed path cannot certify are
    recomputed by the FP64 path (FP64 Gram and solver) in the same call:
    their outputs equal the plain FP64 run's bit for bit, and every other
    pixel agrees with the oracle to the tolerances.  place: the mixed run's
    solver layout; place64: the FP64 layout, which the re-solve follows."""
    rows, cols = 20, 40
    print(f"case resolve n={n} {2 * Nx + 1}x{2 * Ny + 1} {module}/{method} {place} {place64}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=700 + n, patches=6)
    wts, _ = oracle.nmap(slc, Nx, Ny, 0.05)
    out = {}
    for precision in ("fp64", "mixed"):
        with _GenericPlace(n, precision, place if precision == "mixed" else None), \
                _GenericPlace(n, "fp64", place64):
            g, r, res = _run_both(slc, wts, Nx, Ny, module, method, precision, min_nb=5, want_resolved=True)
        out[precision] = (g, res)
    (g64, _), (gmx, res) = out["fp64"], out["mixed"]
    print("re-solved", res.size)
    if n >= 48:   # the corner windows hold fewer samples than dates: |Gamma| is near singular there
        assert res.size > 0
    rr, cc = np.divmod(res, cols)
    for k in range(3):
        a = g64[k][..., rr, cc] if k == 0 else g64[k][rr, cc]
        b = gmx[k][..., rr, cc] if k == 0 else gmx[k][rr, cc]
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
    st = compare_link(gmx, r, r["gap"], stack=slc)
    print(st)
    assert st["untouched_equal"] and st["codes_equal"]
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_link_stbas_and_reference_index(precision):
    n, rows, cols, Nx, Ny = 12, 14, 16, 3, 2
    print(f"case stbas {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=104)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    g, r = _run_both(slc, wts, Nx, Ny, "evd", "STBAS", precision, bw=4, ref=2)
    st = compare_link(g, r, r["gap"])
    print(st)
    assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL
    assert np.all(g[0][2][r["tcorr"] > 0] == 1 + 0j)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_link_failure_codes_and_untouched(precision):
    # 9 window samples for 10 dates: Gamma is rank deficient, evd MLE -> -2;
    # a cleared centre bit leaves the pixel untouched
    n, rows, cols, Nx, Ny = 10, 8, 9, 1, 1
    print(f"case codes {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=105)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wts[3, 4, :] = 0
    g, r = _run_both(slc, wts, Nx, Ny, "evd", "MLE", precision, want_gap=False)
    assert np.array_equal(r["tcorr"] == 0, g[1] == 0)
    assert np.array_equal(np.where(r["tcorr"] < 0, r["tcorr"], 0), np.where(g[1] < 0, g[1], 0))
    assert not g[0][:, 3, 4].any() and g[2][3, 4] == 0
    # phase_link honours minNeighbors
    g2, r2 = _run_both(slc, wts, Nx, Ny, "phase_link", "MLE", precision, min_nb=10, want_gap=False)
    assert np.all(g2[1] == 0) and np.all(r2["tcorr"] == 0)


@pytest.mark.parametrize("module,method", [("phase_link", "MLE"), ("evd", "EVD")])
def test_mixed_scale_ignores_nonfinite_samples(module, method):
    """An infinite and a NaN sample in a stack of amplitudes far beyond the
    fp16 range: the mixed path's per-date scale comes from the finite
    samples, so exactly the pixels whose window holds a bad sample fail, with
    the plain FP64 path's code, and every other pixel agrees with FP64."""
    n, rows, cols, Nx, Ny = 10, 20, 24, 2, 1
    print(f"case non-finite samples {module}/{method}", flush=True)
    slc = (synthetic.stack(n, rows, cols, seed=130, patches=4) * np.float32(1.0e6)).astype(np.complex64)
    bad_at = ((2, 4, 5), (7, 14, 17))   # (date, row, col)
    slc[2, 4, 5] = complex(np.inf, 0.0)
    slc[7, 14, 17] = complex(np.nan, 1.0)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    out = {}
    for precision in ("fp64", "mixed"):
        ph, tc, co = tile.link(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, module=module, method=method,
                               min_neighbors=2, precision=precision)
        torch.cuda.synchronize()
        out[precision] = (ph.cpu().numpy(), tc.cpu().numpy())
    bad = np.zeros((rows, cols), bool)
    for _, r0, c0 in bad_at:
        bad[max(r0 - Ny, 0):r0 + Ny + 1, max(c0 - Nx, 0):c0 + Nx + 1] = True
    (p64, t64), (p16, t16) = out["fp64"], out["mixed"]
    assert np.array_equal(t64 < 0, bad) and np.array_equal(t16 < 0, bad)
    assert np.array_equal(t64[bad], t16[bad])
    good = ~bad
    dphi = np.abs(np.angle(p16[:, good] * np.conj(p64[:, good]))).max(axis=0)
    # near-degenerate pixels aside, the two paths agree to the usual bounds
    assert np.mean(dphi <= PHASE_TOL) >= 0.9
    assert np.mean(np.abs(t16[good] - t64[good]) <= TCORR_TOL) >= 0.9


def test_mixed_precision_tcorr_and_phase_at_cfg1_scale():
    # N=10, 64x64 crop of the 256^2 EVD workload (all-ones mask, 11x11 window)
    n, rows, cols, Nx, Ny = 10, 64, 64, 5, 5
    print("case cfg1 crop", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=101)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    for precision in ("fp64", "mixed"):
        g, r = _run_both(slc, wts, Nx, Ny, "evd", "EVD", precision)
        st = compare_link(g, r, r["gap"])
        print(precision, st)
        assert st["n_well"] == rows * cols
        assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL


@pytest.mark.parametrize("Nx,Ny", [(3, 2), (10, 5), (12, 7)])
def test_staged_gram_tile_variants(Nx, Ny):
    # the staged fp16 hi/lo Gram picks its tile by window size: 8 x 16 pixels
    # on twelve waves (7 x 5 window), 4 x 16 on eight (21 x 11), one pixel on
    # four (25 x 15); each against the FP64 Gram
    n, rows, cols = 30, 20, 40
    print(f"case staged Gram window {2 * Nx + 1}x{2 * Ny + 1}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=500 + Nx, patches=6)
    wts, _ = oracle.nmap(slc, Nx, Ny, 0.05)
    g64, np64 = tile.coherence(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, precision="fp64", min_neighbors=2)
    g16, np16 = tile.coherence(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, precision="mixed", min_neighbors=2)
    torch.cuda.synchronize()
    g64, np64, g16, np16 = g64.cpu().numpy(), np64.cpu().numpy(), g16.cpu().numpy(), np16.cpu().numpy()
    assert np.array_equal(np64, np16)
    linked = np64 >= 0
    assert linked.sum() > 0.5 * rows * cols
    assert np.abs(g16[linked] - g64[linked]).max() < 2e-5


@pytest.mark.parametrize("n,precision", [(10, "mixed"), (30, "mixed"), (48, "mixed"), (72, "mixed"), (100, "mixed"),
                                         (128, "mixed"), (100, "mixed_bf16")])
def test_coherence_matrices(n, precision):
    # fringe_coherence_run: FP64 Gram against a numpy FP64 brute force of
    # evd.cpp:530-582 (complex<float> products, so 1e-6), the hi/lo MFMA Gram
    # against the FP64 one (fp16 pairs with per-date scaling ~22 bits, bf16
    # pairs ~16 bits); above 64 dates the MFMA Gram walks 32-date block pairs,
    # 100 dates leave a partial last block
    rows, cols, Nx, Ny = 10, 12, 3, 2
    print(f"case coherence n={n} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=110 + n, patches=4)
    wts, _ = oracle.nmap(slc, Nx, Ny, 0.05)
    g64, np64 = tile.coherence(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, precision="fp64", min_neighbors=2)
    g16, np16 = tile.coherence(to_dev(slc, torch), wts_dev(wts, torch), Nx, Ny, precision=precision, min_neighbors=2)
    assert g16.dtype == torch.complex64
    torch.cuda.synchronize()
    g64, np64, g16, np16 = g64.cpu().numpy(), np64.cpu().numpy(), g16.cpu().numpy(), np16.cpu().numpy()
    assert np.array_equal(np64, np16)
    for r in range(rows):
        for c in range(cols):
            k = r * cols + c
            sel = []
            for ii in range(-Ny, Ny + 1):
                for jj in range(-Nx, Nx + 1):
                    f = (ii + Ny) * (2 * Nx + 1) + jj + Nx
                    if 0 <= r + ii < rows and 0 <= c + jj < cols and (wts[r, c, f >> 5] >> (f & 31)) & 1:
                        sel.append(slc[:, r + ii, c + jj])
            if not (wts[r, c, (Ny * (2 * Nx + 1) + Nx) >> 5] >> ((Ny * (2 * Nx + 1) + Nx) & 31)) & 1 or len(sel) < 2:
                assert np64[k] == -1
                continue
            Z = np.array(sel, np.complex128)
            G = Z.T @ Z.conj()
            d = np.sqrt(np.real(np.diag(G)))
            want = G / np.outer(d, d)
            assert np64[k] == len(sel)
            assert np.abs(g64[k] - want).max() < 1e-6
            assert np.abs(g16[k] - g64[k]).max() < (1e-4 if precision == "mixed" else 2e-3)
            assert np.all(np.diag(g16[k]) == 1)
