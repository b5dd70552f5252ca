"""fringe_link_run_noalloc (fixed device scratch, no allocation) against
fringe_link_run on the same inputs: bit-identical outputs on two calls in a
row, the second on a smaller tile (the scratch keeps the first call's
contents), and a tile too large for the scratch refused before any launch."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available
from helpers import to_dev, wts_dev

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

import oracle  # noqa: E402
from fringe_amd import _lib, synthetic, tile  # noqa: E402


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(p, slc, wts):
    n, r, c = p.nbands, p.rows, p.cols
    ph = torch.full((n, r, c), complex(7.0, -3.0), dtype=torch.complex64, device="cuda")
    tc = torch.full((r, c), 5.0, dtype=torch.float32, device="cuda")
    co = torch.full((r, c), complex(2.0, 1.0), dtype=torch.complex64, device="cuda")
    s = torch.cuda.current_stream()
    rc = _lib.lib().fringe_link_run_noalloc(ctypes.byref(p), _ptr(slc), _ptr(wts), _ptr(ph), _ptr(tc), _ptr(co),
                                            ctypes.c_void_p(s.cuda_stream))
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return ph.cpu().numpy(), tc.cpu().numpy(), co.cpu().numpy()


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))
               for x, y in zip(a, b))


@pytest.mark.parametrize("precision,module,method", [("mixed", "phase_link", "MLE"), ("fp64", "evd", "MLE"),
                                                    ("mixed", "evd", "EVD")])
def test_noalloc_equals_link_run(precision, module, method):
    # the first tile needs several rounds in the fixed scratch (fringe_link_run
    # takes it in one); the second call is a smaller tile
    cases = [(30, 1024, 1024, 7, 5, 801), (12, 40, 56, 5, 3, 802)]
    for n, rows, cols, Nx, Ny, seed in cases:
        print(f"case noalloc n={n} {rows}x{cols} {module}/{method} {precision}", flush=True)
        slc = synthetic.stack(n, rows, cols, seed=seed, patches=6)
        ds = to_dev(slc, torch)
        if rows * cols > 1 << 16:
            dw, _ = tile.nmap(ds, Nx, Ny, 0.05)   # the GPU map (bit-exact against the oracle's in other tests)
        else:
            dw = wts_dev(oracle.nmap(slc, Nx, Ny, 0.05)[0], torch)
        p = tile.link_params(rows, cols, n, Nx, Ny, module=module, method=method, min_neighbors=5,
                             precision=precision)
        runner = tile.LinkRunner(p)
        ref = runner.run(ds, dw)
        torch.cuda.synchronize()
        ref = tuple(x.cpu().numpy() for x in ref)
        got = _noalloc(p, ds, dw)
        assert (ref[1] != 0).sum() > 0.5 * rows * cols
        assert _same(ref, got)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_register_rows(precision):
    # 33..128 dates: the register-row solvers and (mixed) their FP64 re-solve,
    # whose areas the fixed scratch bounds; the 72-date tile takes several
    # rounds, the second call is a smaller 48-date tile
    cases = [(72, 128, 128, 7, 5, 811), (48, 40, 72, 5, 3, 812)]
    for n, rows, cols, Nx, Ny, seed in cases:
        print(f"case noalloc rows n={n} {rows}x{cols} {precision}", flush=True)
        slc = synthetic.stack(n, rows, cols, seed=seed, patches=6)
        ds = to_dev(slc, torch)
        dw = wts_dev(oracle.nmap(slc, Nx, Ny, 0.05)[0], torch)
        p = tile.link_params(rows, cols, n, Nx, Ny, module="phase_link", method="MLE", min_neighbors=5,
                             precision=precision)
        runner = tile.LinkRunner(p)
        ref = runner.run(ds, dw)
        torch.cuda.synchronize()
        ref = tuple(x.cpu().numpy() for x in ref)
        got = _noalloc(p, ds, dw)
        assert (ref[1] != 0).sum() > 0.5 * rows * cols
        assert _same(ref, got)


def test_noalloc_refuses_a_tile_beyond_the_scratch():
    p = tile.link_params(2048, 2048, 30, 7, 5, precision="mixed")
    dummy = ctypes.c_void_p(256)
    rc = _lib.lib().fringe_link_run_noalloc(ctypes.byref(p), dummy, dummy, dummy, dummy, dummy, None)
    assert rc == 204
    assert b"256 MiB" in _lib.lib().fringe_last_error()
