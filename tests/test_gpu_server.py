"""GPU tests of the tile server: every result is the same alone and in the
mix, admission under a small budget completes every request, and option
errors are reported per request."""
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from fringe_amd import synthetic, tile  # noqa: E402
from helpers import PHASE_TOL, TCORR_TOL, compare_link  # noqa: E402
from fringe_amd.server import TileServer  # noqa: E402

DEV = torch.device("cuda:0")

# (n, rows, cols, Nx, Ny, module, method, precision, server-side nmap)
CASES = [
    (8, 64, 96, 5, 5, "phase_link", "MLE", "mixed", True),
    (16, 96, 80, 7, 5, "phase_link", "MLE", "mixed", False),
    (30, 120, 128, 7, 5, "phase_link", "MLE", "mixed", True),
    (30, 80, 64, 7, 5, "phase_link", "MLE", "fp64", False),
    (12, 72, 72, 5, 3, "evd", "EVD", "mixed", False),
    (40, 48, 56, 4, 3, "phase_link", "MLE", "mixed", True),
]


def _inputs():
    out = []
    for i, (n, rows, cols, Nx, Ny, module, method, prec, srv_nmap) in enumerate(CASES):
        slc = synthetic.stack_torch(n, rows, cols, 300 + i, device=DEV, patches=16)
        wts = None
        if not srv_nmap:
            wts, _ = tile.nmap(slc, Nx, Ny, 0.05)
        p = tile.link_params(rows, cols, n, Nx, Ny, module=module, method=method, min_neighbors=5,
                             precision=prec)
        out.append((slc, wts, p))
    torch.cuda.synchronize()
    return out


def _fetch_all(srv, ids):
    """Wait, fetch and release each request at once (other requests may be
    queued behind it and take its arena space), then read the copies."""
    res = []
    for rid in ids:
        assert srv.wait(rid) == 0
        r = srv.fetch(rid)
        srv.release(rid)
        res.append(r)
    return [[x.cpu() if x is not None else None for x in r] for r in res]


def test_results_equal_alone_and_in_the_mix():
    print("case server alone vs mix", flush=True)
    inp = _inputs()
    with TileServer(8 << 30, streams=4) as srv:
        alone = []
        for slc, wts, p in inp:
            alone += _fetch_all(srv, [srv.submit(slc, p, wts=wts)])
        ids = [srv.submit(slc, p, wts=wts) for slc, wts, p in inp]
        mixed = _fetch_all(srv, ids)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        for x, y in zip(a, m):
            assert (x is None) == (y is None)
            if x is not None:
                assert torch.equal(x, y), f"case {CASES[k]}"
    # and the same as the direct tile API
    for (slc, wts, p), a in zip(inp, alone):
        if wts is None:
            continue
        ph, tc, co = tile.link(slc, wts, p.Nx, p.Ny,
                               module="evd" if p.module == 0 else "phase_link",
                               method={0: "MLE", 1: "EVD", 2: "STBAS"}[p.method], min_neighbors=5,
                               precision={0: "fp64", 1: "mixed", 2: "mixed_bf16"}[p.precision])
        torch.cuda.synchronize()
        assert torch.equal(ph.cpu(), a[0]) and torch.equal(tc.cpu(), a[1]) and torch.equal(co.cpu(), a[2])


def test_server_nmap_matches_tile_nmap():
    print("case server nmap", flush=True)
    n, rows, cols, Nx, Ny = 10, 40, 48, 5, 3
    slc = synthetic.stack_torch(n, rows, cols, 321, device=DEV, patches=8)
    w_ref, c_ref = tile.nmap(slc, Nx, Ny, 0.05)
    p = tile.link_params(rows, cols, n, Nx, Ny, precision="mixed")
    with TileServer(1 << 30, streams=2) as srv:
        rid = srv.submit(slc, p, prob=0.05)
        assert srv.wait(rid) == 0
        ph, tc, co, w, c = srv.fetch(rid)
        torch.cuda.synchronize()
        assert torch.equal(w.cpu(), w_ref.cpu()) and torch.equal(c.cpu(), c_ref.cpu())
        srv.release(rid)


def test_admission_under_a_small_budget():
    print("case server budget", flush=True)
    inp = _inputs()
    with TileServer(1 << 30, streams=4) as probe:
        need = [probe.request_bytes(p, with_mask=w is not None) for _, w, p in inp]
    budget = int(max(need) * 1.5)          # at most one or two requests resident at once
    with TileServer(budget, streams=4) as srv:
        ids = [srv.submit(slc, p, wts=w) for slc, w, p in inp]
        res = _fetch_all(srv, ids)
    with TileServer(8 << 30, streams=4) as srv:
        ref = _fetch_all(srv, [srv.submit(slc, p, wts=w) for slc, w, p in inp])
    for a, b in zip(res, ref):
        for x, y in zip(a, b):
            if x is not None:
                assert torch.equal(x, y)


def test_per_request_errors():
    print("case server errors", flush=True)
    n, rows, cols = 12, 16, 16
    slc = synthetic.stack_torch(n, rows, cols, 5, device=DEV)
    with TileServer(1 << 28, streams=1) as srv:
        bad = tile.link_params(rows, cols, n, 2, 2, module="evd", method="STBAS", bandwidth=0)
        with pytest.raises(RuntimeError, match="code 101"):
            srv.submit(slc, bad, prob=0.05)
        big = tile.link_params(rows, cols, n, 2, 2)
        big.rows = 4096
        big.cols = 4096
        with pytest.raises(ValueError):
            srv.submit(slc, big, prob=0.05)
        huge = torch.zeros((n, 1024, 1024), dtype=torch.complex64, device=DEV)
        p = tile.link_params(1024, 1024, n, 2, 2)
        with pytest.raises(RuntimeError, match="code 204"):
            srv.submit(huge, p, prob=0.05)
        ok = tile.link_params(rows, cols, n, 2, 2)
        rid = srv.submit(slc, ok, prob=0.05)
        assert srv.wait(rid) == 0
        st, code = srv.poll(rid)
        assert st == 2 and code == 0
        q, t = srv.times(rid)
        assert 0 <= q <= t
        srv.release(rid)


def test_destroy_ends_requests_that_can_never_be_admitted():
    print("case server destroy with unreleased outputs", flush=True)
    n, rows, cols, Nx, Ny = 10, 64, 64, 3, 2
    slc = synthetic.stack_torch(n, rows, cols, 9, device=DEV)
    p = tile.link_params(rows, cols, n, Nx, Ny, precision="mixed")
    with TileServer(1 << 30, streams=2) as probe:
        need = probe.request_bytes(p, with_mask=False)
    # room for one request; once it is done its outputs alone leave too
    # little for another
    srv = TileServer(need + 4096, streams=2)
    first = srv.submit(slc, p, prob=0.05)
    assert srv.wait(first) == 0
    # the first request's outputs stay in the arena (never released), so the
    # second cannot be admitted; closing the server must still return
    second = srv.submit(slc, p, prob=0.05)
    st, _ = srv.poll(second)
    assert st == 0
    srv.close()


def test_large_n_requests_against_the_oracle():
    """Requests of the serving mix at 48 and 96 dates, 21 x 11 and 13 x 11
    windows, mixed-precision EMI: each result equals the oracle's to the
    phase and coherence bounds, and is the same alone and in the mix."""
    cases = [(48, 24, 32, 10, 5, True), (96, 20, 24, 6, 5, False), (30, 40, 48, 7, 5, True)]
    inp = []
    for i, (n, rows, cols, Nx, Ny, srv_nmap) in enumerate(cases):
        slc = synthetic.stack(n, rows, cols, seed=900 + i, patches=0 if n == 96 else 8)
        wts = None if srv_nmap else synthetic.full_mask(rows, cols, Nx, Ny)
        p = tile.link_params(rows, cols, n, Nx, Ny, module="phase_link", method="MLE", min_neighbors=5,
                             precision="mixed")
        inp.append((slc, wts, p))
    d = [(torch.from_numpy(slc).to(DEV), None if w is None else torch.from_numpy(w.view(np.int32)).to(DEV), p)
         for slc, w, p in inp]
    print("case server n=48/96/30 against the oracle", flush=True)
    with TileServer(8 << 30, streams=4) as srv:
        alone = []
        for slc, wts, p in d:
            alone += _fetch_all(srv, [srv.submit(slc, p, wts=wts)])
        mixed = _fetch_all(srv, [srv.submit(slc, p, wts=wts) for slc, wts, p in d])
    for a, m in zip(alone, mixed):
        for x, y in zip(a, m):
            if x is not None:
                assert torch.equal(x, y)
    for (slc, wts, p), res, (n, rows, cols, Nx, Ny, srv_nmap) in zip(inp, alone, cases):
        if srv_nmap:
            w_ref, c_ref = oracle.nmap(slc, Nx, Ny, 0.05)
            assert np.array_equal(res[3].numpy().view(np.uint32), w_ref)
            wts = w_ref
        ref = oracle.link(slc, wts, Nx, Ny, module="phase_link", method="MLE", min_neighbors=5, want_gap=True)
        st = compare_link((res[0].numpy(), res[1].numpy(), res[2].numpy()), ref, ref["gap"], stack=slc)
        print(f"  n={n}: {st}")
        assert st["untouched_equal"] and st["codes_equal"]
        assert st["n_well"] >= 0.25 * rows * cols
        assert st["phase_max"] <= PHASE_TOL and st["tcorr_max"] <= TCORR_TOL and st["comp_err"] <= PHASE_TOL
