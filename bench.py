#!/usr/bin/env python3
"""Headline benchmark: pixels/s phase-linked (EMI, N=30), mixed precision vs
the plain FP64 GPU path, with per-tile latency percentiles.

Workload (one "step" = one pass of the phase-linking hot path over the whole
tile, inputs already resident in HBM):
  * N = 30 complex64 SLCs, 2048 x 2048 pixels, half-window 7 x 5 (15 x 11);
  * synthetic circular-Gaussian stack (coherence 0.6 exp(-dt/24 d) + 0.2,
    12-day spacing, 0..3 rad phase ramp), amplitudes replaced by
    sqrt(Gamma(k, theta)) drawn per 32 x 32 patch (k ~ U[1,4], theta
    log-uniform over a decade);
  * SHP mask from this library's KS2 nmap (alpha = 0.05), then
    phase_link EMI with minNeighbors = 5 (phase_link.py default).

Prints one JSON line.  ``--gpus`` above 1 is refused: this build runs on one
GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_HBM = 8000.0          # GB/s, MI355X HBM3E
PEAK_F16_MFMA = 2500.0     # TFLOP/s, dense fp16 MFMA
PEAK_F32_VALU = 157.3      # TFLOP/s, fp32
PROFILE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def gram_flops(n, keff):
    """Useful complex-Gram flops: 8 per complex MAC, both triangles."""
    return 8.0 * n * n * keff


def solve_flops(n):
    """Algorithmic flops of one EMI pixel: inverse of the real SPD |Gamma|
    (N^3), Householder reduction of the Hermitian H to tridiagonal form
    (16/3 N^3), back-transform of one eigenvector (8N^2), Hadamard product
    (2N^2), temporal coherence over N(N-1)/2 pairs (~12 flops each)."""
    return (19.0 / 3.0) * n ** 3 + 16.0 * n * n


def pack_bytes(n, npx):
    return 16.0 * n * npx


def _roof(kernel, bound, achieved, peak, unit, ms, calls, note):
    return {"kernel": kernel, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "ms_per_run": ms, "launches_per_run": calls, "note": note}


def _traffic(kernel, rows, cols, n, Nx, Ny, launches):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the
    counter reads for a device copy of known size); None when absent."""
    path = os.path.join(PROFILE_DIR, "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    key = f"{kernel}:n{n}:{rows}x{cols}:{2 * Nx + 1}x{2 * Ny + 1}"
    e = t.get("kernels", {}).get(key)
    return None if e is None else e.get("bytes_per_launch")


def copy_probe(mib=1024, reps=5):
    """A plain device-to-device copy of known size, run under rocprofv3 --pmc
    to calibrate FETCH_SIZE / WRITE_SIZE against true bytes."""
    import torch
    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": x.numel() * 4, "reps": reps}))


def serve_load(dev, Ns, sides, n_req, rel_rates, seed, budget, n96_side=None, log=None):
    """Serving under load: a pool of n_req EMI requests (N from Ns, sides from
    sides, half-windows 5..10 x 5, random SHP fill 20-100 %) through the tile
    server in mixed precision.  First every request at once (capacity, in
    requests/s), then Poisson arrivals at each fraction of that capacity in
    rel_rates; per point, p50/p99 of submit-to-done latency and the achieved
    pixels/s."""
    import torch
    from fringe_amd import synthetic, tile
    from fringe_amd.server import TileServer
    rng = np.random.default_rng(seed)
    reqs, stacks = [], {}
    for i in range(n_req):
        n = int(rng.choice(Ns))
        side = int(n96_side) if (n > 64 and n96_side) else int(rng.choice(sides))
        Nx, fill = int(rng.integers(5, 11)), float(rng.uniform(0.2, 1.0))
        if (n, side) not in stacks:
            stacks[(n, side)] = synthetic.stack_torch(n, side, side, 1000 + len(stacks), device=dev)
        wts = synthetic.random_mask_torch(side, side, Nx, 5, fill, 2000 + i, device=dev)
        reqs.append((n, side, Nx, fill, wts))
    torch.cuda.synchronize()

    def params(n, side, Nx):
        return tile.link_params(side, side, n, Nx, 5, module="phase_link", method="MLE", min_neighbors=5,
                                precision="mixed")

    def run(arrivals):
        with TileServer(budget, streams=4) as srv:
            t0 = time.perf_counter()
            ids = []
            for (n, side, Nx, fill, wts), at in zip(reqs, arrivals):
                now = time.perf_counter() - t0
                if at > now:
                    time.sleep(at - now)
                ids.append(srv.submit(stacks[(n, side)], params(n, side, Nx), wts=wts))
            lat = []
            for rid in ids:
                assert srv.wait(rid) == 0
                lat.append(srv.times(rid)[1])
            span = time.perf_counter() - t0
            linked = 0
            for rid in ids:
                linked += int((srv.fetch(rid)[1] != 0).sum())
                srv.release(rid)
        return np.array(lat), span, linked

    run(np.zeros(len(reqs)))   # warm-up: kernel loading, allocator
    lat, span, linked = run(np.zeros(len(reqs)))
    cap = len(reqs) / span
    out = {"requests": len(reqs), "capacity_req_per_s": cap, "saturated_pixels_per_s": linked / span,
           "saturated_p50_ms": float(np.percentile(lat, 50)), "saturated_p99_ms": float(np.percentile(lat, 99)),
           "sweep": []}
    for rel in rel_rates:
        arr = np.cumsum(np.random.default_rng(seed + int(rel * 1000)).exponential(1.0 / (rel * cap), size=len(reqs)))
        lat, span, linked = run(arr)
        pt = {"offered_frac_of_capacity": rel, "offered_req_per_s": rel * cap, "p50_ms": float(np.percentile(lat, 50)),
              "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()), "pixels_per_s": linked / span}
        out["sweep"].append(pt)
        if log:
            print(f"serving at {rel:.2f} x capacity: p50 {pt['p50_ms']:.1f} ms p99 {pt['p99_ms']:.1f} ms",
                  file=log, flush=True)
    out["mix"] = [[n, s, nx, round(f, 2)] for n, s, nx, f, _ in reqs]
    return out


def batch(args):
    """Mixed-N batch: a pool of EMI tiles (N in {8, 16, 30, 48, 96}, sides
    256/512 (N = 96: 256), half-windows 5..10 x 5, fill 20-100 %) linked by one
    fringe_link_batch_run call (one persistent solve launch per date class and
    round) against the same tiles linked one after another by
    fringe_link_run; outputs compared bit for bit."""
    import torch
    from fringe_amd import synthetic, tile
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(106)
    tiles = []
    for i in range(args.batch_tiles):
        n = int(rng.choice([8, 16, 30, 48, 96]))
        side = 256 if n > 64 else int(rng.choice([256, 512]))
        Nx, fill = int(rng.integers(5, 11)), float(rng.uniform(0.2, 1.0))
        slc = synthetic.stack_torch(n, side, side, 3000 + i, device=dev, patches=8)
        w = synthetic.random_mask_torch(side, side, Nx, 5, fill, 3100 + i, device=dev)
        tiles.append((slc, w, tile.link_params(side, side, n, Nx, 5, min_neighbors=5, precision="mixed")))
    runners = [tile.LinkRunner(p, device=dev) for _, _, p in tiles]
    b = tile.BatchRunner(tiles, device=dev)

    def one_by_one():
        for r, (slc, w, _) in zip(runners, tiles):
            r.run(slc, w)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    t_sep = timed(one_by_one, args.steps)
    t_bat = timed(b.run, args.steps)
    def bits(x):
        return (torch.view_as_real(x) if x.is_complex() else x).contiguous().view(torch.int32)

    same = all(torch.equal(bits(x), bits(y)) for r, o in zip(runners, b.outs)
               for x, y in zip((r.phase, r.tcorr, r.comp), o))
    px = sum(int((o[1] != 0).sum()) for o in b.outs)
    print(json.dumps({
        "metric": "mixed-N batch: one fringe_link_batch_run against fringe_link_run tile by tile (EMI, mixed)",
        "value": px / (t_bat / 1e3), "unit": "pixels/s", "n_gpus": 1, "higher_is_better": True,
        "batch_ms": t_bat, "one_by_one_ms": t_sep, "speedup": t_sep / t_bat, "identical_outputs": bool(same),
        "tiles": [[p.nbands, p.rows, p.Nx] for _, _, p in tiles], "steps": args.steps,
        "data": "synthetic (seeded torch generator on the GPU; random SHP masks of the drawn fill)"}))


def serve(args):
    """Serving workload (cfg5): EMI requests with N in {8, 16, 30, 48, 96},
    sides 256..2048 (N = 96: 512), half-windows 5x5 to 10x5 and random SHP
    fill 20-100 %, through the tile server in mixed precision under a
    device-memory budget: capacity, then Poisson arrivals at 0.5, 0.8 and 1.2
    times it (p50/p99 latency); then the same requests one after another on
    the plain FP64 path."""
    import torch
    from fringe_amd import synthetic, tile
    dev = torch.device("cuda:0")
    free, _ = torch.cuda.mem_get_info()
    budget = int(min(args.budget_gb * 1e9, free - 16e9))
    load = serve_load(dev, [8, 16, 30, 48, 96], [256, 512, 1024, 2048], args.serve_requests, [0.5, 0.8, 1.2], 105,
                      budget, n96_side=512, log=sys.stderr)
    # plain FP64, serial, same requests (the mix is rebuilt from the same seed)
    rng = np.random.default_rng(105)
    stacks = {}
    t64, linked64 = 0.0, 0
    for i, (n, side, Nx, fill) in enumerate(load["mix"]):
        if (n, side) not in stacks:
            stacks[(n, side)] = synthetic.stack_torch(n, side, side, 1000 + len(stacks), device=dev)
        wts = synthetic.random_mask_torch(side, side, Nx, 5, fill, 2000 + i, device=dev)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        r = tile.LinkRunner(tile.link_params(side, side, n, Nx, 5, module="phase_link", method="MLE",
                                             min_neighbors=5, precision="fp64"), device=dev)
        r.run(stacks[(n, side)], wts)
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t2
        linked64 += int((r.tcorr != 0).sum())
        del r
        print(f"fp64 request {i}: N={n} side={side}", file=sys.stderr, flush=True)
    below = load["sweep"][0]
    out = {
        "metric": "serving: p50/p99 tile latency and aggregate pixels/s (EMI, mixed) vs serial plain FP64",
        "value": load["saturated_pixels_per_s"], "unit": "pixels/s", "n_gpus": 1, "higher_is_better": True,
        "p50_ms": below["p50_ms"], "p99_ms": below["p99_ms"], "offered_frac_of_capacity": below["offered_frac_of_capacity"],
        "serial_fp64_pixels_per_s": linked64 / t64, "serial_fp64_s": t64,
        "speedup_vs_serial_fp64": load["saturated_pixels_per_s"] / (linked64 / t64), "budget_bytes": budget,
        "serving": load,
        "data": "synthetic (seeded; random SHP masks of the drawn fill)",
        "config": {"workload": "N in {8,16,30,48,96}, sides {256,512,1024,2048} (N=96: 512), Nx in 5..10, Ny=5, "
                               "fill U[0.2,1], phase_link EMI minNeighbors=5, 4 server streams"},
    }
    print(json.dumps(out))


def cfg3(args):
    """Covariance-dominated case: N = 64 dates, 4096 columns, half-window
    10 x 5 (231 neighbours, all-ones mask).  On a band of output rows (with its
    Ny halo), the per-pixel Frobenius error of Gamma from the fp16 and bf16
    hi/lo MFMA Grams against the FP64 Gram, and the phase error of the EVD
    top eigenvector (evd module, EVD) against the FP64 path."""
    import torch
    from fringe_amd import synthetic, tile
    dev = torch.device("cuda:0")
    n, cols, Nx, Ny = 64, 4096, 10, 5
    out_rows = args.cfg3_rows
    rows = out_rows + 2 * Ny
    torch.cuda.manual_seed(103)
    slc = synthetic.stack_torch(n, rows, cols, 103, device=dev)
    W = (2 * Nx + 1) * (2 * Ny + 1)
    words = np.zeros((W + 31) // 32, np.uint32)
    for f in range(W):
        words[f >> 5] |= np.uint32(1 << (f & 31))
    wts = torch.from_numpy(np.broadcast_to(words.view(np.int32), (rows, cols, len(words))).copy()).to(dev)
    rb, re_ = Ny, Ny + out_rows

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) * 1e3

    (g64, np64), t64 = timed(lambda: tile.coherence(slc, wts, Nx, Ny, precision="fp64", module="evd",
                                                    row_begin=rb, row_end=re_))
    ok = np64 >= 0
    nrm = torch.linalg.matrix_norm(g64[ok]).double()
    res = {}
    for prec in ("mixed", "mixed_bf16"):
        (g, npx_), tg = timed(lambda: tile.coherence(slc, wts, Nx, Ny, precision=prec, module="evd",
                                                     row_begin=rb, row_end=re_))
        err = torch.linalg.matrix_norm(g[ok].to(torch.complex128) - g64[ok]).double()
        rel = err / nrm
        res[prec] = {"gamma_frob_err_p50": float(err.median()), "gamma_frob_err_p99": float(torch.quantile(err, 0.99)),
                     "gamma_frob_err_max": float(err.max()), "gamma_frob_rel_max": float(rel.max()),
                     "gram_ms": tg}
        del g
    del g64
    torch.cuda.empty_cache()
    phases = {}
    for prec in ("fp64", "mixed", "mixed_bf16"):
        p = tile.link_params(rows, cols, n, Nx, Ny, module="evd", method="EVD", precision=prec,
                             row_begin=rb, row_end=re_)
        r = tile.LinkRunner(p, device=dev)
        r.run(slc, wts)
        torch.cuda.synchronize()
        (_, tl) = timed(lambda: r.run(slc, wts))
        phases[prec] = (r.phase[:, rb:re_].clone(), r.tcorr[rb:re_].clone(), tl)
        del r
    ph64, tc64, t_link64 = phases["fp64"]
    okp = tc64 > 0
    for prec in ("mixed", "mixed_bf16"):
        ph, tc, tl = phases[prec]
        e = torch.angle(ph * torch.conj(ph64)).abs().amax(dim=0)[okp].double()
        res[prec].update({"evd_phase_err_p99_rad": float(torch.quantile(e.float(), 0.99)),
                          "evd_phase_err_max_rad": float(e.max()),
                          "frac_phase_err_gt_1e-3": float((e > 1e-3).double().mean()),
                          "tcorr_err_max": float((tc - tc64).abs()[okp].max()), "link_ms": tl})
    out = {"metric": "cfg3: Gamma Frobenius error and EVD phase error of hi/lo MFMA Grams vs FP64 Gram (N=64)",
           "config": {"workload": f"N={n}, {out_rows} output rows x {cols} cols (+{Ny}-row halos) of a 4096^2 "
                                  f"tile, window {2 * Nx + 1}x{2 * Ny + 1}, all-ones mask, evd EVD",
                      "pixels": int(ok.sum())},
           "data": "synthetic (seeded, generated on the GPU)", "fp64_gram_ms": t64, "fp64_link_ms": t_link64,
           "fp16_hilo": res["mixed"], "bf16_hilo": res["mixed_bf16"]}
    print(json.dumps(out))


def cfg1(args):
    """The reference's CPU-sized case: N = 10 dates of 256 x 256, phase ramp
    phi_n(col) = n * 3 rad * col / (cols - 1), 11 x 11 all-ones window, evd
    EVD.  GPU FP64 and mixed link times on the full tile, the oracle (CPU,
    reference semantics, FP64 LAPACK) timed on the same full tile as the CPU
    baseline, the GPU phases against the oracle's, and the RMSE of the
    recovered ramp."""
    import torch
    from fringe_amd import synthetic, tile
    dev = torch.device("cuda:0")
    n, rows, cols, Nx, Ny = 10, 256, 256, 5, 5
    slc_np = synthetic.stack(n, rows, cols, seed=101)
    wts_np = synthetic.full_mask(rows, cols, Nx, Ny)
    slc = torch.from_numpy(slc_np).to(dev)
    wts = torch.from_numpy(wts_np.view(np.int32)).to(dev)
    res, runs = {}, {}
    for prec in ("fp64", "mixed"):
        r = tile.LinkRunner(tile.link_params(rows, cols, n, Nx, Ny, module="evd", method="EVD", precision=prec),
                            device=dev)
        for _ in range(args.warmup):
            r.run(slc, wts)
        torch.cuda.synchronize()
        t = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            r.run(slc, wts)
            torch.cuda.synchronize()
            t.append((time.perf_counter() - t0) * 1e3)
        runs[prec] = (r.phase.cpu().numpy(), r.tcorr.cpu().numpy())
        res[prec] = {"ms_median": float(np.median(t)), "pixels_per_s": rows * cols / (np.median(t) / 1e3)}
        del r
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref = oracle.link(slc_np, wts_np, Nx, Ny, module="evd", method="EVD", want_gap=True, threads=threads)
    dt = time.perf_counter() - t0
    well = (ref["tcorr"] > 0) & (ref["gap"] >= 1e-3)
    truth = np.exp(1j * synthetic.ramp_phase(n, cols))[:, None, :]   # referenced to date 0 (zero phase)
    for prec, (ph, tc) in runs.items():
        e = np.abs(np.angle(ph * np.conj(ref["phase"]))).max(axis=0)
        res[prec].update({
            "phase_err_vs_oracle_max_rad": float(e[well].max()),
            "tcorr_err_vs_oracle_max": float(np.abs(tc - ref["tcorr"])[well].max()),
            "ramp_rmse_deg": float(np.degrees(np.sqrt(np.mean(np.angle(ph * np.conj(truth)) ** 2))))})
    out = {"metric": "cfg1: pixels/s linked (evd EVD, N=10, 256^2, 11x11 all-ones window), GPU vs the CPU oracle",
           "config": {"workload": f"N={n}, {rows}x{cols} px, window {2 * Nx + 1}x{2 * Ny + 1}, all-ones mask, "
                                  "phase ramp 0..3 rad per date step, evd EVD", "pixels": rows * cols},
           "data": "synthetic (seeded numpy generator, seed 101)",
           "fp64": res["fp64"], "mixed": res["mixed"],
           "ramp_rmse_deg_oracle": float(np.degrees(np.sqrt(np.mean(np.angle(ref["phase"] * np.conj(truth)) ** 2)))),
           "cpu_baseline": {"value": rows * cols / dt, "unit": "pixels/s", "cores": threads, "kind": "port",
                            "sample": f"the full {rows}x{cols} tile ({dt:.2f} s), oracle/fringe_oracle.cpp, "
                                      "FP64 LAPACK, OpenMP"}}
    print(json.dumps(out))


def e2e(args):
    """End-to-end time of the file-based modules on the headline shape (N = 30
    dates, rows x cols, 15 x 11 window, gamma-patch amplitudes): the stack is
    written as per-date ENVI files with raw VRTs and a stack VRT (the
    tops2vrt layout), the nmap module writes the SHP mask, and the phase_link
    module (EMI, minNeighbors 5, the reference's default block memory)
    reads stack and mask, links and writes one ENVI raster per date plus
    tcorr and the compressed SLC.  Also the host-buffer rate: the stack and
    mask in host memory, copied to the GPU, linked, and the outputs copied
    back (PCIe-inclusive, no file I/O)."""
    import datetime
    import shutil
    import tempfile
    import torch
    from fringe_amd import raster, synthetic, tile
    from fringe_amd.cli import nmap as nmap_cli, phase_link as pl_cli
    dev = torch.device("cuda:0")
    n, rows, cols, Nx, Ny = args.n, args.rows, args.cols, args.nx, args.ny
    torch.cuda.manual_seed(args.seed)
    slc = synthetic.stack_torch(n, rows, cols, args.seed, device=dev, patches=32).cpu().numpy()
    root = tempfile.mkdtemp(prefix="fringe_e2e_")
    out = {"metric": "end-to-end phase_link module time (files in, files out) and host-buffer rate",
           "config": {"workload": f"phase_link EMI, N={n}, {rows}x{cols} px, window {2 * Nx + 1}x{2 * Ny + 1}, "
                                  "gamma-patch amplitudes, nmap KS2 mask, minNeighbors=5", "pixels": rows * cols},
           "data": "synthetic (seeded), written to local disk"}
    try:
        t = time.perf_counter()
        srcs, d0 = [], datetime.date(2022, 1, 1)
        os.makedirs(os.path.join(root, "slcs"))
        for k in range(n):
            d = (d0 + datetime.timedelta(12 * k)).strftime("%Y%m%d")
            os.makedirs(os.path.join(root, "SLC", d))
            f = os.path.join(root, "SLC", d, d + ".slc.full")
            slc[k].tofile(f)
            v = os.path.join(root, "slcs", d + ".vrt")
            with open(v, "w") as fh:
                fh.write(raster.raw_vrt(f, cols, rows))
            srcs.append((v, d, True))
        base = os.path.join(root, "slcs_base.vrt")
        with open(base, "w") as fh:
            fh.write(raster.stack_vrt(srcs, cols, rows))
        out["write_stack_s"] = time.perf_counter() - t
        print(f"stack written in {out['write_stack_s']:.1f} s", file=sys.stderr, flush=True)
        wf, cf = os.path.join(root, "nmap", "wts"), os.path.join(root, "nmap", "count")
        t = time.perf_counter()
        assert nmap_cli.main(["-i", base, "-o", wf, "-c", cf, "-x", str(Nx), "-y", str(Ny)]) == 0
        out["nmap_module_s"] = time.perf_counter() - t
        print(f"nmap module {out['nmap_module_s']:.1f} s", file=sys.stderr, flush=True)
        for prec in ("mixed", "fp64"):
            od = os.path.join(root, "pl_" + prec)
            t = time.perf_counter()
            assert pl_cli.main(["-i", base, "-w", wf, "-o", od, "-x", str(Nx), "-y", str(Ny), "-n", "5",
                                "--precision", prec]) == 0
            dt = time.perf_counter() - t
            out[f"phase_link_module_{prec}_s"] = dt
            out[f"phase_link_module_{prec}_pixels_per_s"] = rows * cols / dt
            print(f"phase_link module ({prec}) {dt:.1f} s", file=sys.stderr, flush=True)
            shutil.rmtree(od)
        # host buffers: H2D + link + D2H, no file I/O
        nul = (((2 * Nx + 1) * (2 * Ny + 1)) + 31) // 32
        wts = np.stack([raster.read(wf, b + 1, dtype=np.uint32) for b in range(nul)], axis=-1)
        slc_h = torch.from_numpy(slc).pin_memory()
        wts_h = torch.from_numpy(wts.view(np.int32)).pin_memory()
        r = tile.LinkRunner(tile.link_params(rows, cols, n, Nx, Ny, module="phase_link", method="MLE",
                                             min_neighbors=5, precision="mixed"), device=dev)
        ph_h = torch.empty((n, rows, cols), dtype=torch.complex64).pin_memory()
        tc_h = torch.empty((rows, cols), dtype=torch.float32).pin_memory()
        co_h = torch.empty((rows, cols), dtype=torch.complex64).pin_memory()
        times = []
        for it in range(args.warmup + 3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            sd, wd = slc_h.to(dev, non_blocking=True), wts_h.to(dev, non_blocking=True)
            ph, tc, co = r.run(sd, wd)
            ph_h.copy_(ph, non_blocking=True)
            tc_h.copy_(tc, non_blocking=True)
            co_h.copy_(co, non_blocking=True)
            torch.cuda.synchronize()
            if it >= args.warmup:
                times.append(time.perf_counter() - t)
        out["host_buffers_mixed_s"] = float(np.median(times))
        out["host_buffers_mixed_pixels_per_s"] = rows * cols / float(np.median(times))
        out["bytes_over_pcie"] = int(slc.nbytes + wts.nbytes + ph_h.numel() * 8 + tc_h.numel() * 4 + co_h.numel() * 8)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    print(json.dumps(out))


def cfg4(args):
    """Sequential estimator on device-resident data (sequential.py:144-255):
    120 dates of 1024 x 1024 (12-day spacing, same coherence model), split into
    mini-stacks of lengths drawn from 8..20; mini-stack k = compressed SLCs of
    mini-stacks 1..k-1 + its own dates, linked by the evd module (MLE) with
    reference index k-1; its compressed SLC joins the next ones; the stack of
    compressed SLCs is linked last (datum connection).  SHP mask: nmap KS2
    over all 120 dates, half-window 7 x 5.  Timed for the mixed and the plain
    FP64 path."""
    import torch
    from fringe_amd import synthetic, tile
    dev = torch.device("cuda:0")
    n_all, side, Nx, Ny = 120, 1024, 7, 5
    rng = np.random.default_rng(104)
    sizes = []
    while sum(sizes) < n_all:
        sizes.append(int(rng.integers(8, 21)))
    sizes[-1] -= sum(sizes) - n_all
    if sizes[-1] < 1:
        sizes.pop()
    torch.cuda.manual_seed(104)
    slc = synthetic.stack_torch(n_all, side, side, 104, device=dev, patches=32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wts, _ = tile.nmap(slc, Nx, Ny, 0.05)
    torch.cuda.synchronize()
    t_nmap = time.perf_counter() - t0
    print(f"nmap over {n_all} dates: {t_nmap:.2f} s", file=sys.stderr, flush=True)

    def run(prec):
        comps, outs, per = [], [], []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        s0 = 0
        for k, L in enumerate(sizes, start=1):
            stack = torch.cat([torch.stack(comps), slc[s0:s0 + L]]) if comps else slc[s0:s0 + L].contiguous()
            t1 = time.perf_counter()
            ph, tc, co = tile.link(stack, wts, Nx, Ny, module="evd", method="MLE", ref_index=k - 1,
                                   precision=prec)
            torch.cuda.synchronize()
            per.append((stack.shape[0], (time.perf_counter() - t1) * 1e3))
            comps.append(co)
            outs.append(ph[k - 1:])
            s0 += L
        t1 = time.perf_counter()
        ph, tc, co = tile.link(torch.stack(comps), wts, Nx, Ny, module="evd", method="MLE", ref_index=0,
                               precision=prec)
        torch.cuda.synchronize()
        per.append((len(comps), (time.perf_counter() - t1) * 1e3))
        total = time.perf_counter() - t_start
        return total, per, outs, ph

    res = {}
    for prec in ("mixed", "fp64"):
        run(prec)   # warm-up (kernel loading for every N)
        res[prec] = run(prec)
        print(f"{prec}: {res[prec][0]:.2f} s", file=sys.stderr, flush=True)
    err = 0.0
    for a, b in zip(res["mixed"][2] + [res["mixed"][3]], res["fp64"][2] + [res["fp64"][3]]):
        ok = (b.abs() > 0).all(dim=0)
        err = max(err, float(torch.angle(a * torch.conj(b)).abs().amax(dim=0)[ok].quantile(0.999)))
    links = len(sizes) + 1
    out = {"metric": "cfg4: sequential mini-stacks (pixels/s over all mini-stack and datum links)",
           "value": links * side * side / res["mixed"][0], "unit": "pixels/s",
           "fp64_value": links * side * side / res["fp64"][0], "speedup_vs_fp64": res["fp64"][0] / res["mixed"][0],
           "mini_stack_sizes": sizes, "links": links, "mixed_s": res["mixed"][0], "fp64_s": res["fp64"][0],
           "per_link_ms_mixed": [[nk, round(ms, 2)] for nk, ms in res["mixed"][1]],
           "per_link_ms_fp64": [[nk, round(ms, 2)] for nk, ms in res["fp64"][1]],
           "phase_err_p999_rad_vs_fp64": err, "nmap_s": t_nmap,
           "config": {"workload": f"{n_all} dates, {side}x{side}, mini-stacks U{{8..20}}, evd MLE, window "
                                  f"{2 * Nx + 1}x{2 * Ny + 1}, nmap KS2 over all dates"},
           "data": "synthetic (seeded, generated on the GPU)"}
    print(json.dumps(out))


def vs_oracle(oracle, res, band, wb, rr, Nx, Ny, min_nb, threads, gpu):
    """Phase, temporal-coherence and code agreement of the GPU outputs with
    the oracle's on the oracle's row band.  A pixel off by more than 1e-3 rad
    is counted as ill-posed when its relative eigen-gap (recomputed by the
    oracle for the rows that hold such pixels) is below 1e-3."""
    import torch
    r0, r1, a0 = rr
    ph_o = res["phase"][:, r0 - a0:r1 - a0]
    tc_o = res["tcorr"][r0 - a0:r1 - a0]
    out = {"reference": "oracle (FP64 LAPACK restatement of phase_link.cpp:479-666)",
           "rows": [r0, r1], "pixels": int(tc_o.size)}
    gaps = {}
    for name, (ph, tcg) in gpu.items():
        ph_g = ph[:, r0:r1].cpu().numpy()
        tc_g = tcg[r0:r1].cpu().numpy()
        ok = (tc_o > 0) & (tc_g > 0)
        err = np.abs(np.angle(ph_g * np.conj(ph_o))).max(axis=0)
        bad = ok & (err > 1e-3)
        for row in np.unique(np.nonzero(bad)[0]):
            if row not in gaps:
                g = oracle.link(band, wb, Nx, Ny, module="phase_link", method="MLE", min_neighbors=min_nb,
                                rows_range=(r0 - a0 + row, r0 - a0 + row + 1), want_gap=True, threads=threads)
                gaps[row] = g["gap"][r0 - a0 + row]
        n_well_bad = sum(int((gaps[row][np.nonzero(bad[row])[0]] >= 1e-3).sum())
                         for row in np.unique(np.nonzero(bad)[0]))
        out[name] = {"phase_err_max_rad": float(err[ok].max()) if ok.any() else 0.0,
                     "phase_err_p99_rad": float(np.percentile(err[ok], 99)) if ok.any() else 0.0,
                     "n_phase_err_gt_1e-3": int(bad.sum()), "n_gt_1e-3_with_gap_ge_1e-3": n_well_bad,
                     "tcorr_err_max": float(np.abs(tc_g - tc_o)[ok].max()) if ok.any() else 0.0,
                     "codes_equal": bool(np.array_equal(np.where(tc_o <= 0, tc_o, 1), np.where(tc_g <= 0, tc_g, 1)))}
    return out


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=2048)
    ap.add_argument("--cols", type=int, default=2048)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--nx", type=int, default=7)
    ap.add_argument("--ny", type=int, default=5)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=512, help="output rows of the CPU oracle sample (about 10 s on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--device-gen", action="store_true",
                    help="generate the tile with torch's generator on the GPU instead of numpy PCG64 on the host")
    ap.add_argument("--seed", type=int, default=102)
    ap.add_argument("--copy-probe", action="store_true", help="run only the copy kernel used to calibrate PMC bytes")
    ap.add_argument("--serve", action="store_true", help="serving workload: Poisson stream of mixed tiles")
    ap.add_argument("--e2e", action="store_true",
                    help="file-based nmap + phase_link modules on the headline shape, and the host-buffer rate")
    ap.add_argument("--cfg1", action="store_true",
                    help="N=10, 256^2, evd EVD: GPU FP64 and mixed against the CPU oracle on the full tile")
    ap.add_argument("--cfg3", action="store_true",
                    help="N=64, 21x11 window: coherence and EVD phase errors of the hi/lo Grams against FP64")
    ap.add_argument("--cfg3-rows", type=int, default=32, help="output rows of the N=64 sample (4096 columns)")
    ap.add_argument("--cfg4", action="store_true",
                    help="sequential mini-stacks of uneven length (120 dates, 1024^2), on device-resident data")
    ap.add_argument("--batch", action="store_true",
                    help="mixed-N batch of tiles in one call against the same tiles one by one")
    ap.add_argument("--batch-tiles", type=int, default=24)
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-under-load sample of the headline run")
    ap.add_argument("--load-requests", type=int, default=40,
                    help="requests per point of the headline run's serving-under-load sample")
    ap.add_argument("--rate", type=float, default=40.0, help="mean arrivals per second")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.serve:
        serve(args)
        return
    if args.cfg1:
        cfg1(args)
        return
    if args.e2e:
        e2e(args)
        return
    if args.cfg3:
        cfg3(args)
        return
    if args.cfg4:
        cfg4(args)
        return
    if args.batch:
        batch(args)
        return
    import torch
    from fringe_amd import _lib, synthetic, tile

    L = _lib.lib()
    n, rows, cols, Nx, Ny = args.n, args.rows, args.cols, args.nx, args.ny
    dev = torch.device("cuda:0")
    torch.cuda.manual_seed(args.seed)
    if args.device_gen:
        slc = synthetic.stack_torch(n, rows, cols, args.seed, device=dev, patches=32)
        data = "synthetic (seeded torch generator on the GPU)"
    else:
        # numpy PCG64(seed): the same tile can be rebuilt on any host
        slc = torch.from_numpy(synthetic.stack(n, rows, cols, args.seed, patches=32)).to(dev)
        data = f"synthetic (numpy PCG64 seed {args.seed}, fringe_amd.synthetic.stack, copied to the GPU)"
    wts, cnt = tile.nmap(slc, Nx, Ny, 0.05)
    torch.cuda.synchronize()
    L.fringe_timing_reset()
    L.fringe_timing_enable(1)
    wts, cnt = tile.nmap(slc, Nx, Ny, 0.05)
    torch.cuda.synchronize()
    L.fringe_timing_enable(0)
    nm_per = {k: v for k, v in _lib.stage_times().items() if v[1]}
    min_nb = 5
    W = (2 * Nx + 1) * (2 * Ny + 1)
    cnt_f = cnt.to(torch.float64)
    fill = float((cnt_f / W).mean())

    def make(precision):
        return tile.LinkRunner(tile.link_params(rows, cols, n, Nx, Ny, module="phase_link", method="MLE",
                                                min_neighbors=min_nb, precision=precision), device=dev)

    def timed(runner, steps, warmup):
        for _ in range(warmup):
            runner.run(slc, wts)
        torch.cuda.synchronize()
        L.fringe_timing_reset()
        L.fringe_timing_enable(1)
        lat = []
        for _ in range(steps):
            t0 = time.perf_counter()
            runner.run(slc, wts)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
        L.fringe_timing_enable(0)
        return np.array(lat), _lib.stage_times()

    mixed = make("mixed")
    lat, stages = timed(mixed, args.steps, args.warmup)
    n_resolved = int(mixed.resolved_pixels().size)
    tc = mixed.tcorr
    linked = int((tc != 0).sum())
    ms = float(lat.mean())
    value = linked / (ms / 1e3)
    ph_mixed = mixed.phase.clone()

    fp64 = make("fp64")
    lat64, stages64 = timed(fp64, args.fp64_steps, 1)
    ms64 = float(lat64.mean())
    value64 = int((fp64.tcorr != 0).sum()) / (ms64 / 1e3)
    ok = (fp64.tcorr > 0) & (mixed.tcorr > 0)
    err = torch.angle(ph_mixed * torch.conj(fp64.phase)).abs().amax(dim=0)[ok]
    tcerr = (mixed.tcorr - fp64.tcorr).abs()[ok]
    acc = {
        "phase_err_max_rad": float(err.max()), "phase_err_p99_rad": float(torch.quantile(err.float()[:1 << 24], 0.99)),
        "frac_phase_err_gt_1e-3": float((err > 1e-3).double().mean()),
        "tcorr_err_max": float(tcerr.max()), "reference": "FP64 GPU path (itself checked against the FP64 oracle in tests/)",
    }

    # ---- roofline of every stage, the dominant one on the line -----------
    steps_timed = args.steps
    per = {k: (v[0] / max(steps_timed, 1), v[1] / max(steps_timed, 1)) for k, v in stages.items() if v[1]}
    keff = float(cnt_f[(tc != 0)].sum())
    rl = {}
    if "gram" in per:
        g_ms, g_calls = per["gram"]
        rl["gram"] = _roof("gram_mfma", "mfma", gram_flops(n, keff) / (g_ms / 1e3) / 1e12, PEAK_F16_MFMA,
                           "TFLOP/s", g_ms, g_calls, "8*K_eff*N^2 flops per linked pixel (hi/lo passes not counted)")
        try:   # rocprofv3 MFMA counters of this workload (profiles/pmc_mfma_gram_*.md)
            with open(os.path.join(PROFILE_DIR, "pmc_mfma.json")) as f:
                pm = json.load(f)
            rl["gram"]["rocprof_mfma_busy_frac"] = pm["mfma_busy_frac"]
            rl["gram"]["rocprof_mfma_executed_frac_of_peak"] = pm["frac_of_f16_peak_executed"]
        except (OSError, ValueError, KeyError):
            pass
    if "solve" in per:
        s_ms, s_calls = per["solve"]
        rl["solve"] = _roof("solve_reg_kernel", "valu", solve_flops(n) * linked / (s_ms / 1e3) / 1e12, PEAK_F32_VALU,
                            "TFLOP/s", s_ms, s_calls,
                            "(19/3)N^3 + 16N^2 flops per linked pixel: |Gamma| inverse, Hermitian tridiagonal "
                            "reduction, back-transform, Hadamard product, temporal coherence")
    if "pack" in per:
        p_ms, p_calls = per["pack"]
        rl["pack"] = _roof("pack_hilo", "hbm", pack_bytes(n, rows * cols) / (p_ms / 1e3) / 1e9, PEAK_HBM,
                           "GB/s", p_ms, p_calls, "16N bytes per pixel of the tile (stack read twice, records written)")
    if "nmap_prep" in nm_per:
        m_ms, m_calls = nm_per["nmap_prep"]
        rl["nmap_prep"] = _roof("nmap_prep", "hbm", rows * cols * 12 * n / (m_ms / 1e3) / 1e9, PEAK_HBM, "GB/s",
                                m_ms, m_calls, "12N bytes per pixel (complex64 read, sorted amplitudes written)")
    dom = max(("gram", "solve", "pack"), key=lambda k: per.get(k, (0, 0))[0])
    roof = {k: v for k, v in rl[dom].items() if k in ("kernel", "bound", "achieved", "peak", "unit", "frac", "note")}
    roof["traffic"] = _traffic(roof["kernel"], rows, cols, n, Nx, Ny, per[dom][1])

    # ---- CPU baseline: the oracle on a bounded row band -------------------
    cpu = None
    acc_oracle = None
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
            import oracle
            threads = min(16, os.cpu_count() or 1)
            r0 = rows // 2
            r1 = min(rows, r0 + args.cpu_rows)
            a0, a1 = max(0, r0 - Ny), min(rows, r1 + Ny)
            band = slc[:, a0:a1].cpu().numpy()
            wb = wts[a0:a1].cpu().numpy().view(np.uint32)
            t0 = time.perf_counter()
            res = oracle.link(band, wb, Nx, Ny, module="phase_link", method="MLE", min_neighbors=min_nb,
                              rows_range=(r0 - a0, r1 - a0), threads=threads)
            dt = time.perf_counter() - t0
            lk = int((res["tcorr"][r0 - a0:r1 - a0] != 0).sum())
            cpu = {"value": lk / dt, "unit": "pixels/s", "cores": threads, "kind": "port",
                   "sample": f"{r1 - r0} output rows x {cols} cols of the same tile ({lk} pixels linked, {dt:.1f} s), "
                             "oracle/fringe_oracle.cpp with FP64 LAPACK, OpenMP"}
            acc_oracle = vs_oracle(oracle, res, band, wb, (r0, r1, a0), Nx, Ny, min_nb, threads,
                                   {"mixed": (ph_mixed, tc), "fp64": (fp64.phase, fp64.tcorr)})
        except Exception as e:  # the baseline is reported, never the product
            cpu = {"error": str(e)}

    # ---- serving under load: p50/p99 below capacity and a short sweep -----
    load = None
    if not args.no_serve:
        free, _ = torch.cuda.mem_get_info()
        load = serve_load(dev, [8, 16, 30, 48, 96], [256, 512], args.load_requests, [0.5, 0.8, 1.2], 105,
                          int(min(32e9, free - 8e9)), n96_side=256, log=sys.stderr)
        load["config"] = ("tile server, 4 streams, mixed-precision EMI requests: N in {8,16,30,48,96}, sides "
                          "256/512 (N=96: 256), Nx 5..10, Ny 5, fill U[0.2,1]; Poisson arrivals at fractions of "
                          "the measured capacity")

    out = {
        "metric": "pixels/s phase-linked (EMI, N=30) mixed vs FP64 on GPU; p99 tile latency",
        "value": value, "unit": "pixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "mixed: f16 hi/lo MFMA Gram (f32 acc), f32 eigensolver, f64 |Gamma| inverse; "
                 "pixels whose f32 eigenpair is not certified to 5e-4 rad re-solved in f64 in the same call",
        "data": data,
        "config": {"workload": f"phase_link EMI, N={n}, {rows}x{cols} px, window {2*Nx+1}x{2*Ny+1}, "
                               f"nmap KS2 mask (alpha=0.05) on gamma-patch amplitudes, minNeighbors={min_nb}",
                   "pixels_linked": linked, "shp_fill_mean": fill},
        "fp64_value": value64, "fp64_ms_per_step": ms64, "speedup_vs_fp64": value / value64,
        "p50_tile_ms": float(np.percentile(lat, 50)), "p99_tile_ms": float(np.percentile(lat, 99)),
        "p50_tile_ms_under_load": load["sweep"][0]["p50_ms"] if load else None,
        "p99_tile_ms_under_load": load["sweep"][0]["p99_ms"] if load else None,
        "serving_under_load": load,
        "resolved_fp64_pixels": n_resolved, "resolved_fp64_frac": n_resolved / max(linked, 1),
        "accuracy_vs_fp64": acc, "accuracy_vs_oracle": acc_oracle,
        "stages_ms_per_step": {k: v[0] for k, v in per.items()},
        "nmap_ms": {k: v[0] for k, v in nm_per.items()},
        "stages_fp64_ms_per_step": {k: v[0] / max(args.fp64_steps, 1) for k, v in stages64.items() if v[1]},
        "roofline": roof, "roofline_stages": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
