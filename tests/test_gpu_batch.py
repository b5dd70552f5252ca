"""Batched tiles and persistent solve launches: one persistent launch per
date class drains the pixels of several tiles of different date counts,
windows and SHP fills (tiles on the register-row solver above 64 dates get a
grid launch each); the outputs equal each tile's own fringe_link_run bit for
bit.  Also: a single tile under a persistent solver schedule equals the
grid launch, and every Gram variant the schedule table can force gives the
same coherence matrices."""
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


def _bits(x):
    return np.ascontiguousarray(x.cpu().numpy()).view(np.uint32)


def _same(a, b):
    return all(np.array_equal(_bits(x), _bits(y)) for x, y in zip(a, b))


def _tiles(specs, module, method, precision):
    out = []
    for k, (n, rows, cols, Nx, Ny, fill) in enumerate(specs):
        slc = synthetic.stack(n, rows, cols, seed=600 + k, patches=6)
        ds = to_dev(slc, torch)
        if fill is None:
            w = wts_dev(oracle.nmap(slc, Nx, Ny, 0.05)[0], torch)
        else:
            w = synthetic.random_mask_torch(rows, cols, Nx, Ny, fill, 700 + k, device="cuda")
        p = tile.link_params(rows, cols, n, Nx, Ny, module=module, method=method, min_neighbors=5,
                             precision=precision, ref_index=min(2, n - 1))
        out.append((ds, w, p))
    return out


# (n, rows, cols, Nx, Ny, fill); fill None = the SHP map of the stack
# (100 dates: the MFMA Gram's complex64 coherence into the class of up to
# 128 dates; 136: the FP64 Gram's complex128 into the class above)
SPECS = [(10, 24, 40, 5, 3, None), (24, 20, 36, 7, 5, 0.4), (30, 18, 44, 7, 5, None), (17, 16, 30, 3, 2, 0.9),
         (40, 14, 32, 5, 4, None), (64, 12, 26, 6, 4, 0.7), (80, 12, 22, 7, 5, None), (100, 10, 18, 5, 4, None),
         (136, 10, 16, 5, 4, 0.8)]


class _Sched:
    """Sets one schedule key for the duration of a block, restoring it."""

    def __init__(self, kind, key, value):
        self.L, self.kind, self.key, self.value = _lib.lib(), kind, key, value

    def __enter__(self):
        self.old = self.L.fringe_schedule_get(self.kind, *self.key)
        if self.value is not None:
            assert self.L.fringe_schedule_set(self.kind, *self.key, self.value) == 0

    def __exit__(self, *exc):
        self.L.fringe_schedule_set(self.kind, *self.key, self.old)


@pytest.mark.parametrize("precision,module,method,place", [
    ("mixed", "phase_link", "MLE", None), ("fp64", "phase_link", "MLE", None), ("mixed", "evd", "MLE", None),
    ("mixed", "evd", "EVD", None), ("mixed", "phase_link", "MLE", "lds")])
def test_batch_equals_per_tile_runs(precision, module, method, place):
    """place: the 80-date tile's solver (FRINGE_SCHED_GENERIC_A) -- rows in
    registers by default in the mixed precisions, one grid launch per tile;
    with the LDS work matrix it joins the persistent launch of its class."""
    print(f"case batch {len(SPECS)} tiles {module}/{method} {precision} {place}", flush=True)
    tiles = _tiles(SPECS, module, method, precision)
    v = None if place is None else _lib.GENERIC_A_PLACES[place]
    with _Sched(_lib.FRINGE_SCHED_GENERIC_A, (80, 0 if precision == "fp64" else 1, 0), v):
        ref = []
        for ds, w, p in tiles:
            r = tile.LinkRunner(p)
            ref.append(tuple(x.clone() for x in r.run(ds, w)))
        torch.cuda.synchronize()
        b = tile.BatchRunner(tiles)
        got = b.run()
        torch.cuda.synchronize()
    for k, (r, g) in enumerate(zip(ref, got)):
        assert (r[1] != 0).sum() > 0
        assert _same(r, g), f"tile {k}"


def test_batch_across_rounds():
    """Three 512 x 256 tiles at 30 dates hold more coherence matrices than one
    ~2 GB round: the rounds split tiles and mix them, with the same outputs."""
    specs = [(30, 512, 256, 7, 5, 0.8), (30, 512, 256, 7, 5, 0.5), (12, 512, 256, 5, 5, 0.9)]
    print("case batch across rounds", flush=True)
    tiles = []
    for k, (n, rows, cols, Nx, Ny, fill) in enumerate(specs):
        ds = synthetic.stack_torch(n, rows, cols, 610 + k, device="cuda", patches=8)
        w = synthetic.random_mask_torch(rows, cols, Nx, Ny, fill, 710 + k, device="cuda")
        tiles.append((ds, w, tile.link_params(rows, cols, n, Nx, Ny, min_neighbors=5, precision="mixed")))
    ref = [tuple(x.clone() for x in tile.LinkRunner(p).run(ds, w)) for ds, w, p in tiles]
    b = tile.BatchRunner(tiles)
    got = b.run()
    torch.cuda.synchronize()
    for r, g in zip(ref, got):
        assert _same(r, g)


@pytest.mark.parametrize("n,precision", [(30, "mixed"), (48, "mixed"), (20, "fp64"), (12, "mixed"), (16, "fp64"),
                                         (27, "mixed"), (30, "fp64"),
                                         (80, "mixed"), (100, "mixed"), (136, "mixed"), (100, "fp64")])
def test_persistent_schedule_equals_grid(n, precision):
    # up to 32 dates the grid launch's tridiagonal reductions process 16, 24,
    # 28, 30 or 32 columns (the fewest that hold n) and the persistent
    # launch's 32: the same bits
    L = _lib.lib()
    rows, cols, Nx, Ny = 40, 64, 7, 5
    print(f"case persistent schedule n={n} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=620 + n, patches=6)
    ds = to_dev(slc, torch)
    w = synthetic.random_mask_torch(rows, cols, Nx, Ny, 0.7, 720 + n, device="cuda")
    p = tile.link_params(rows, cols, n, Nx, Ny, min_neighbors=5, precision=precision)
    key = (n, 0, 0 if precision == "fp64" else 1)
    old = L.fringe_schedule_get(_lib.FRINGE_SCHED_SOLVE, *key)
    # above 32 dates the persistent launch serves the one-row-per-lane solver
    # (up to 64 dates) and the LDS work matrix solver, not the register rows
    place = _Sched(_lib.FRINGE_SCHED_GENERIC_A, (n, key[2], 0), 0 if n > 32 else None)
    place.__enter__()
    try:
        assert L.fringe_schedule_set(_lib.FRINGE_SCHED_SOLVE, *key, 0) == 0
        grid = tuple(x.clone() for x in tile.LinkRunner(p).run(ds, w))
        for waves, grab in ((16, 4), (8, 1), (32, 16)):
            assert L.fringe_schedule_set(_lib.FRINGE_SCHED_SOLVE, *key, (waves << 8) | grab) == 0
            pers = tile.LinkRunner(p).run(ds, w)
            torch.cuda.synchronize()
            assert _same(grid, pers)
    finally:
        L.fringe_schedule_set(_lib.FRINGE_SCHED_SOLVE, *key, old)
        place.__exit__()


@pytest.mark.parametrize("Nx,Ny", [(7, 5), (10, 5), (3, 2)])
def test_gram_variants_agree(Nx, Ny):
    L = _lib.lib()
    n, rows, cols = 30, 24, 48
    print(f"case Gram variants {2 * Nx + 1}x{2 * Ny + 1}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=630 + Nx, patches=6)
    ds = to_dev(slc, torch)
    w = wts_dev(oracle.nmap(slc, Nx, Ny, 0.05)[0], torch)
    old = L.fringe_schedule_get(_lib.FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny)
    res = {}
    try:
        for name, v in _lib.GRAM_VARIANTS.items():
            assert L.fringe_schedule_set(_lib.FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny, v) == 0
            g, npx = tile.coherence(ds, w, Nx, Ny, precision="mixed", min_neighbors=2)
            torch.cuda.synchronize()
            res[name] = (g.cpu().numpy(), npx.cpu().numpy())
    finally:
        L.fringe_schedule_set(_lib.FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny, old)
    g0, n0 = res["default"]
    for name, (g, npx) in res.items():
        assert np.array_equal(npx, n0), name
        linked = n0 >= 0
        assert np.abs(g[linked] - g0[linked]).max() < 1e-6, name


@pytest.mark.parametrize("n,precision", [(80, "mixed"), (96, "mixed"), (72, "fp64")])
def test_work_matrix_placement_agrees(n, precision):
    """Above 64 dates the solver's work matrix in LDS or in global memory (the
    FRINGE_SCHED_GENERIC_A schedule) gives the same outputs bit for bit."""
    L = _lib.lib()
    rows, cols, Nx, Ny = 12, 24, 6, 5
    print(f"case work matrix placement n={n} {precision}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=640 + n)
    ds = to_dev(slc, torch)
    w = synthetic.random_mask_torch(rows, cols, Nx, Ny, 0.9, 740 + n, device="cuda")
    p = tile.link_params(rows, cols, n, Nx, Ny, min_neighbors=5, precision=precision)
    pk = 0 if precision == "fp64" else 1
    old = L.fringe_schedule_get(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0)
    out = []
    try:
        for v in (0, 1):
            assert L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0, v) == 0
            out.append(tuple(x.clone() for x in tile.LinkRunner(p).run(ds, w)))
            torch.cuda.synchronize()
    finally:
        L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0, old)
    assert (out[0][1] != 0).sum() > 0 and _same(out[0], out[1])

