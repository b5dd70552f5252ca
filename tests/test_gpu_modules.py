"""GPU end-to-end tests of the module drop-in: on-disk stacks in the layout
tops2vrt.py writes, run through the nmap / evd / phase_link / sequential
drivers, checked against the CPU oracle."""
import glob
import os

import numpy as np
import pytest

from conftest import gpu_available
from helpers import PHASE_TOL, TCORR_TOL, wrapped_err

pytestmark = pytest.mark.gpu
if not gpu_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import oracle  # noqa: E402
from fringe_amd import raster, synthetic  # noqa: E402
from fringe_amd.cli import evd as evd_cli  # noqa: E402
from fringe_amd.cli import nmap as nmap_cli  # noqa: E402
from fringe_amd.cli import phase_link as pl_cli  # noqa: E402


def write_stack(root, slc, start=20220101, amp_const=None):
    """SLC/<date>/<date>.slc.full + raw VRTs + slcs_base.vrt (tops2vrt.py:144-225)."""
    import datetime
    n, rows, cols = slc.shape
    t0 = datetime.date(2022, 1, 1)
    srcs, meta, dates = [], [], []
    os.makedirs(os.path.join(root, "slcs"), exist_ok=True)
    for t in range(n):
        d = (t0 + datetime.timedelta(12 * t)).strftime("%Y%m%d")
        dd = os.path.join(root, "SLC", d)
        os.makedirs(dd, exist_ok=True)
        f = os.path.join(dd, d + ".slc.full")
        slc[t].tofile(f)
        v = os.path.join(root, "slcs", d + ".vrt")
        open(v, "w").write(raster.raw_vrt(f, cols, rows))
        srcs.append((v, d, True))
        meta.append({} if amp_const is None else {"amplitudeConstant": amp_const[t]})
        dates.append(d)
    base = os.path.join(root, "slcs_base.vrt")
    open(base, "w").write(raster.stack_vrt(srcs, cols, rows, metadata=meta))
    return base, dates


def read_outputs(folder, dates, rows, cols):
    ph = np.stack([raster.read(os.path.join(folder, d + ".slc")) for d in dates])
    tc = raster.read(os.path.join(folder, "tcorr.bin"), dtype=np.float32)
    co = raster.read(os.path.join(folder, "compslc.bin"))
    return ph, tc, co


def hdr_fields(path):
    """key -> value of an ENVI header (<path>.hdr, the SUFFIX=ADD name)."""
    out = {}
    for line in open(path + ".hdr"):
        if "=" in line:
            k, v = line.split("=", 1)
            out[k.strip().lower()] = v.strip()
    return out


def test_nmap_and_phase_link_modules(tmp_path):
    n, rows, cols, Nx, Ny = 10, 28, 33, 3, 2
    slc = synthetic.stack(n, rows, cols, seed=31, patches=6)
    base, dates = write_stack(str(tmp_path), slc)
    wf, cf = str(tmp_path / "KS2" / "nmap"), str(tmp_path / "KS2" / "count")
    print("case: nmap module", flush=True)
    assert nmap_cli.main(["-i", base, "-o", wf, "-c", cf, "-x", str(Nx), "-y", str(Ny)]) == 0
    nul = oracle.mask_words(Nx, Ny)
    wts = np.stack([raster.read(wf, b + 1, dtype=np.uint32) for b in range(nul)], axis=-1)
    cnt = raster.read(cf, dtype=np.float32).astype(np.int32)
    wr, cr = oracle.nmap(slc, Nx, Ny, 0.05)
    assert np.array_equal(wts, wr) and np.array_equal(cnt, cr)
    assert raster.metadata(wf, "HALFWINDOWX", "ENVI") == str(Nx)
    # headers as GDAL's ENVI driver writes them (nmap.cpp:236-253, :541-548):
    # UInt32 BIP mask with the half-window keys, Int16 count
    h = hdr_fields(wf)
    assert (h["samples"], h["lines"], h["bands"]) == (str(cols), str(rows), str(nul))
    assert (h["data type"], h["interleave"], h["byte order"]) == ("13", "bip", "0")
    assert (h["halfwindowx"], h["halfwindowy"]) == (str(Nx), str(Ny))
    h = hdr_fields(cf)
    assert (h["samples"], h["lines"], h["bands"], h["data type"]) == (str(cols), str(rows), "1", "2")
    print("case: phase_link module, one block and many blocks", flush=True)
    out1, out2 = str(tmp_path / "pl1"), str(tmp_path / "pl2")
    assert pl_cli.main(["-i", base, "-w", wf, "-o", out1, "-x", str(Nx), "-y", str(Ny)]) == 0
    # tiny memsize/blocksize: many row blocks with Ny-row overlap
    assert pl_cli.main(["-i", base, "-w", wf, "-o", out2, "-x", str(Nx), "-y", str(Ny), "-l", "8", "-r", "1"]) == 0
    g1 = read_outputs(out1, dates, rows, cols)
    g2 = read_outputs(out2, dates, rows, cols)
    for a, b in zip(g1, g2):
        assert np.array_equal(a, b)
    ref = oracle.link(slc, wr, Nx, Ny, module="phase_link", method="MLE", min_neighbors=5)
    ok = ref["tcorr"] > 0
    assert np.array_equal(ok, g1[1] > 0)
    assert wrapped_err(g1[0], ref["phase"])[:, ok].max() <= PHASE_TOL
    assert np.abs(g1[1] - ref["tcorr"])[ok].max() <= TCORR_TOL
    assert os.path.exists(os.path.join(out1, dates[0] + ".slc.vrt"))
    # output headers (evd.cpp:266-366): CFloat32 phases and compressed SLC, Float32 tcorr
    for name, code in ((dates[0] + ".slc", "6"), ("tcorr.bin", "4"), ("compslc.bin", "6")):
        h = hdr_fields(os.path.join(out1, name))
        assert (h["samples"], h["lines"], h["bands"], h["data type"]) == (str(cols), str(rows), "1", code)
        assert h["byte order"] == "0"


@pytest.mark.parametrize("method", ["EVD", "MLE"])
def test_evd_module(tmp_path, method):
    n, rows, cols, Nx, Ny = 8, 20, 21, 2, 2
    slc = synthetic.stack(n, rows, cols, seed=41)
    base, dates = write_stack(str(tmp_path), slc)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": Nx, "HALFWINDOWY": Ny})
    out = str(tmp_path / "evd")
    print(f"case: evd module {method}", flush=True)
    assert evd_cli.main(["-i", base, "-w", wf, "-o", out, "-x", str(Nx), "-y", str(Ny), "-m", method,
                         "--precision", "fp64"]) == 0
    ph, tc, co = read_outputs(out, dates, rows, cols)
    ref = oracle.link(slc, wts, Nx, Ny, module="evd", method=method)
    ok = ref["tcorr"] > 0
    assert np.array_equal(np.where(ref["tcorr"] <= 0, ref["tcorr"], 1), np.where(tc <= 0, tc, 1))
    assert wrapped_err(ph, ref["phase"])[:, ok].max() <= PHASE_TOL
    assert np.abs(tc - ref["tcorr"])[ok].max() <= TCORR_TOL


def test_sequential_against_oracle_replay(tmp_path):
    from fringe_amd.sequential import run_sequential
    n, rows, cols, Nx, Ny, L = 12, 16, 18, 3, 2, 5
    slc = synthetic.stack(n, rows, cols, seed=51)
    # SLC/<date>/<date>.slc as ENVI (Stack.py globs */*.slc)
    import datetime
    t0 = datetime.date(2022, 1, 1)
    dates = []
    for t in range(n):
        d = (t0 + datetime.timedelta(12 * t)).strftime("%Y%m%d")
        os.makedirs(tmp_path / "SLC" / d)
        raster.write_envi(str(tmp_path / "SLC" / d / (d + ".slc")), slc[t])
        dates.append(d)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": Nx, "HALFWINDOWY": Ny})
    print("case: sequential 12 dates, mini-stacks of 5", flush=True)
    res = run_sequential(str(tmp_path / "SLC"), wf, str(tmp_path / "seq"), miniStackSize=L,
                         halfWindowX=Nx, halfWindowY=Ny, minNeighbors=5, verbose=False)
    assert all(rc == 0 for _, rc in res) and len(res) == 4
    # oracle replay of the scheme (sequential.py:193-254, evd module MLE)
    comps = []
    k = 0
    for s0 in range(0, n, L):
        k += 1
        s1 = min(n, s0 + L)
        stack = np.concatenate([np.array(comps).reshape(-1, rows, cols), slc[s0:s1]]) if comps else slc[s0:s1]
        ref = oracle.link(stack.astype(np.complex64), wts, Nx, Ny, module="evd", method="MLE",
                          mini_stack_count=k)
        folder = str(tmp_path / "seq" / "miniStacks" / f"{dates[s0]}_{dates[s1 - 1]}" / "EVD")
        comp_dates = sorted(os.path.basename(os.path.dirname(p)) for p in
                            glob.glob(str(tmp_path / "seq" / "compressedSlc" / "*" / "*.slc")))
        band_dates = [d for d in comp_dates if d < dates[s0]] + dates[s0:s1]
        ph = np.stack([raster.read(os.path.join(folder, d + ".slc")) for d in band_dates])
        tc = raster.read(os.path.join(folder, "tcorr.bin"), dtype=np.float32)
        comp = raster.read(str(tmp_path / "seq" / "compressedSlc" / dates[s1 - 1] / (dates[s1 - 1] + ".slc")))
        ok = ref["tcorr"] > 0
        assert ok.sum() > 0.8 * rows * cols
        assert wrapped_err(ph, ref["phase"])[:, ok].max() <= PHASE_TOL
        assert np.abs(tc - ref["tcorr"])[ok].max() <= TCORR_TOL
        rel = np.abs(comp - ref["comp"])[ok] / np.maximum(np.abs(ref["comp"][ok]), 1e-6)
        assert rel.max() < 1e-3
        # the next mini-stack (and its replay) sees this run's compressed SLC
        comps.append(comp)
    # datum connection, replayed on the same compressed SLCs
    dc = str(tmp_path / "seq" / "Datum_connection" / "EVD")
    ref = oracle.link(np.array(comps).astype(np.complex64), wts, Nx, Ny, module="evd", method="MLE")
    ph = np.stack([raster.read(os.path.join(dc, dates[min(n, s0 + L) - 1] + ".slc")) for s0 in range(0, n, L)])
    ok = ref["tcorr"] > 0
    assert wrapped_err(ph, ref["phase"])[:, ok].max() <= PHASE_TOL


def test_sequential_uneven_ministacks_120_dates(tmp_path):
    """120 dates split into mini-stacks of lengths drawn from 8..20 (the
    uneven-length extension of sequential.py:193-240), every mini-stack and
    the datum connection replayed through the oracle."""
    import datetime
    from fringe_amd.sequential import run_sequential
    n, rows, cols, Nx, Ny = 120, 16, 20, 5, 4
    rng = np.random.default_rng(104)
    sizes = []
    while sum(sizes) < n:
        sizes.append(int(rng.integers(8, 21)))
    slc = synthetic.stack(n, rows, cols, seed=104)
    t0 = datetime.date(2022, 1, 1)
    dates = []
    for t in range(n):
        d = (t0 + datetime.timedelta(12 * t)).strftime("%Y%m%d")
        os.makedirs(tmp_path / "SLC" / d)
        raster.write_envi(str(tmp_path / "SLC" / d / (d + ".slc")), slc[t])
        dates.append(d)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": Nx, "HALFWINDOWY": Ny})
    print(f"case: sequential 120 dates, mini-stacks {sizes}", flush=True)
    res = run_sequential(str(tmp_path / "SLC"), wf, str(tmp_path / "seq"), halfWindowX=Nx, halfWindowY=Ny,
                         minNeighbors=5, verbose=False, mini_stack_sizes=sizes)
    assert all(rc == 0 for _, rc in res) and len(res) == len(sizes) + 1
    comps, ends = [], []
    s0 = 0
    for k, size in enumerate(sizes, start=1):
        s1 = min(n, s0 + size)
        stack = np.concatenate([np.array(comps).reshape(-1, rows, cols), slc[s0:s1]]) if comps else slc[s0:s1]
        assert stack.shape[0] == (k - 1) + (s1 - s0)
        ref = oracle.link(stack.astype(np.complex64), wts, Nx, Ny, module="evd", method="MLE",
                          mini_stack_count=k, want_gap=True)
        folder = str(tmp_path / "seq" / "miniStacks" / f"{dates[s0]}_{dates[s1 - 1]}" / "EVD")
        band_dates = ends + dates[s0:s1]
        ph = np.stack([raster.read(os.path.join(folder, d + ".slc")) for d in band_dates])
        tc = raster.read(os.path.join(folder, "tcorr.bin"), dtype=np.float32)
        comp = raster.read(str(tmp_path / "seq" / "compressedSlc" / dates[s1 - 1] / (dates[s1 - 1] + ".slc")))
        ok = (ref["tcorr"] > 0) & (ref["gap"] >= 1e-3)
        assert ok.sum() > 0.9 * rows * cols
        assert np.array_equal(np.where(ref["tcorr"] <= 0, ref["tcorr"], 1), np.where(tc <= 0, tc, 1))
        assert wrapped_err(ph, ref["phase"])[:, ok].max() <= PHASE_TOL
        assert np.abs(tc - ref["tcorr"])[ok].max() <= TCORR_TOL
        amp = np.abs(stack[k - 1:]).mean(axis=0)
        assert (np.abs(comp - ref["comp"]) / amp)[ok].max() <= PHASE_TOL
        comps.append(comp)   # the next mini-stack (and its replay) sees this run's compressed SLC
        ends.append(dates[s1 - 1])
        s0 = s1
    dc = str(tmp_path / "seq" / "Datum_connection" / "EVD")
    ref = oracle.link(np.array(comps).astype(np.complex64), wts, Nx, Ny, module="evd", method="MLE",
                      want_gap=True)
    ph = np.stack([raster.read(os.path.join(dc, d + ".slc")) for d in ends])
    ok = (ref["tcorr"] > 0) & (ref["gap"] >= 1e-3)
    assert ok.sum() > 0.9 * rows * cols
    assert wrapped_err(ph, ref["phase"])[:, ok].max() <= PHASE_TOL


def test_datum_adjustment_wrapped_and_unwrapped(tmp_path):
    # adjustMiniStacks.py after a sequential run: wrapped phases multiplied by
    # the mini-stack's datum phase as GDAL's "mul" pixel function does it (in
    # double, rounded to CFloat32); unwrapped phases added, with the
    # time-series VRT's ScaleOffset = -phase at the reference pixel
    from fringe_amd.cli import adjust_ministacks
    from fringe_amd.sequential import run_sequential
    import datetime
    import re
    n, rows, cols, Nx, Ny, L = 8, 12, 14, 2, 2, 4
    slc = synthetic.stack(n, rows, cols, seed=52)
    t0 = datetime.date(2022, 1, 1)
    dates = []
    os.makedirs(tmp_path / "vrts")
    for t in range(n):
        d = (t0 + datetime.timedelta(12 * t)).strftime("%Y%m%d")
        os.makedirs(tmp_path / "SLC" / d)
        f = str(tmp_path / "SLC" / d / (d + ".slc"))
        raster.write_envi(f, slc[t])
        open(str(tmp_path / "vrts" / (d + ".vrt")), "w").write(raster.raw_vrt(f, cols, rows))
        dates.append(d)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": Nx, "HALFWINDOWY": Ny})
    print("case: datum adjustment", flush=True)
    res = run_sequential(str(tmp_path / "SLC"), wf, str(tmp_path / "seq"), miniStackSize=L,
                         halfWindowX=Nx, halfWindowY=Ny, minNeighbors=5, verbose=False)
    assert all(rc == 0 for _, rc in res)
    ms, dc = str(tmp_path / "seq" / "miniStacks"), str(tmp_path / "seq" / "Datum_connection")
    out = str(tmp_path / "adjusted")
    assert adjust_ministacks.main(["-s", str(tmp_path / "vrts"), "-m", ms, "-d", dc, "-M", str(L), "-o", out]) == 0
    for i, d in enumerate(dates):
        s0 = (i // L) * L
        last = dates[min(n, s0 + L) - 1]
        a = raster.read(os.path.join(ms, f"{dates[s0]}_{last}", "EVD", d + ".slc"))
        b = raster.read(os.path.join(dc, "EVD", last + ".slc"))
        want = (a.astype(np.complex128) * b.astype(np.complex128)).astype(np.complex64)
        got = raster.read(os.path.join(out, d + ".slc"))
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        assert np.array_equal(raster.read(os.path.join(out, d + ".slc.vrt")), got)
    # unwrapped: synthetic unwrapped phases beside the EVD folders
    rng = np.random.default_rng(3)
    for s0 in range(0, n, L):
        last = dates[min(n, s0 + L) - 1]
        os.makedirs(os.path.join(ms, f"{dates[s0]}_{last}", "unwrap"))
        for d in dates[s0:s0 + L]:
            raster.write_envi(os.path.join(ms, f"{dates[s0]}_{last}", "unwrap", d + ".unw"),
                              rng.normal(size=(rows, cols)).astype(np.float32) * 10)
    os.makedirs(os.path.join(dc, "unwrap"))
    for s0 in range(0, n, L):
        last = dates[min(n, s0 + L) - 1]
        raster.write_envi(os.path.join(dc, "unwrap", last + ".unw"), rng.normal(size=(rows, cols)).astype(np.float32))
    out2 = str(tmp_path / "adjusted_unw")
    ts = str(tmp_path / "ts.vrt")
    assert adjust_ministacks.main(["-s", str(tmp_path / "vrts"), "-m", ms, "-d", dc, "-M", str(L), "-o", out2,
                                   "--unwrapped", "-t", ts, "-x", "3", "-y", "5"]) == 0
    offsets = [float(v) for v in re.findall(r"<ScaleOffset>([^<]+)</ScaleOffset>", open(ts).read())]
    assert len(offsets) == n
    for i, d in enumerate(dates):
        s0 = (i // L) * L
        last = dates[min(n, s0 + L) - 1]
        a = raster.read(os.path.join(ms, f"{dates[s0]}_{last}", "unwrap", d + ".unw"), dtype=np.float32)
        b = raster.read(os.path.join(dc, "unwrap", last + ".unw"), dtype=np.float32)
        got = raster.read(os.path.join(out2, d + ".phase"), dtype=np.float32)
        assert np.array_equal(got, (a.astype(np.float64) + b.astype(np.float64)).astype(np.float32))
        assert offsets[i] == pytest.approx(-float(got[5, 3]), rel=1e-6)


@pytest.mark.parametrize("bands,coh", [([1], False), ([1, 2], False), ([1, 2], True)])
def test_despeck_module_matches_oracle(tmp_path, bands, coh):
    # despeck.py through the GPU module, several row blocks (memsize 1 MB),
    # against the oracle's whole-image sums: bit-exact
    from fringe_amd.cli import despeck as despeck_cli
    n, rows, cols, Nx, Ny = 3, 70, 40, 3, 2
    print(f"case despeck bands={bands} coherence={coh}", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=61, patches=5)
    sv, _ = write_stack(tmp_path, slc)
    wts, _ = oracle.nmap(slc, Nx, Ny, 0.05)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": Nx, "HALFWINDOWY": Ny})
    out = str(tmp_path / "despeck" / "out.bin")
    argv = ["-i", sv, "-o", out, "-w", wf, "-x", str(Nx), "-y", str(Ny), "-r", "1", "-l", "8",
            "-b"] + [str(b) for b in bands] + (["-c"] if coh else [])
    assert despeck_cli.main(argv) == 0
    want = oracle.despeck(slc[bands[0] - 1], wts, Nx, Ny, b2=slc[bands[1] - 1] if len(bands) == 2 else None,
                          coherence=coh)
    if len(bands) == 1:
        got = raster.read(out, dtype=np.float32)
        assert np.array_equal(got, want.real)
    else:
        got = raster.read(out)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_calamp_module_feeds_nmap(tmp_path):
    # calamp.py writes a VRT of the stack whose bands carry amplitudeConstant
    # (6 significant digits); nmap.py reads it back (nmap.cpp:204-233)
    from fringe_amd.cli import calamp as calamp_cli
    n, rows, cols, Nx, Ny = 6, 40, 36, 3, 2
    print("case calamp", flush=True)
    slc = synthetic.stack(n, rows, cols, seed=71, patches=6)
    slc *= np.linspace(0.5, 3.0, n)[:, None, None].astype(np.float32)
    sv, dates = write_stack(tmp_path, slc)
    mask = (np.arange(rows * cols).reshape(rows, cols) % 7 != 0).astype(np.uint8)
    mf = str(tmp_path / "mask")
    raster.write_envi(mf, mask)
    out = str(tmp_path / "calib.vrt")
    assert calamp_cli.main(["-i", sv, "-o", out, "-m", mf, "-r", "1", "-l", "8"]) == 0
    want = oracle.calamp(slc, mask)
    got = [raster.metadata(out, "amplitudeConstant", "slc", band=b + 1) for b in range(n)]
    assert got == [f"{v:g}" for v in want]
    assert [raster.metadata(out, "Date", "slc", band=b + 1) for b in range(n)] == dates
    assert np.array_equal(raster.read(out, band=3), raster.read(sv, band=3))
    ac = np.array([float(v) for v in got])
    wr, cr = oracle.nmap(slc, Nx, Ny, 0.05, amp_const=ac)
    wf, cf = str(tmp_path / "nmap" / "wts"), str(tmp_path / "nmap" / "count")
    assert nmap_cli.main(["-i", out, "-o", wf, "-c", cf, "-x", str(Nx), "-y", str(Ny)]) == 0
    nul = oracle.mask_words(Nx, Ny)
    wg = np.stack([raster.read(wf, band=k + 1, dtype=np.uint32) for k in range(nul)], axis=-1)
    assert np.array_equal(wg, wr)
