"""CPU tests of libfringe_amd.so's host side: exported C ABI, KS decision
table, ENVI/VRT reading and writing, option objects and the module error
codes that are decided before any GPU work."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from fringe_amd import _lib, raster, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fringe_amd.h")).read()
    decl = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(fringe_\w+)\s*\(", hdr, re.M))
    L = _lib.lib()
    missing = [s for s in sorted(decl) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= decl


@pytest.mark.parametrize("n", [8, 10, 16, 20, 28, 30, 48, 64, 96])
def test_ks_table_matches_oracle_decisions(n):
    L = _lib.lib()
    tab = np.zeros(n + 1, np.uint8)
    margin = L.fringe_ks_table(n, 0.05, tab.ctypes.data)
    for k in range(n + 1):
        p = oracle.kolmogorov_prob(k / n * np.sqrt(n * n / (2.0 * n)))
        assert tab[k] == (1 if p >= 0.05 else 0)
    assert margin > 1e-6          # every decision is far from the threshold
    kstar = int(np.argmin(tab)) if not tab.all() else n + 1
    assert np.all(tab[:kstar] == 1) and np.all(tab[kstar:] == 0)


def test_ks_table_thresholds_quoted_for_the_configs():
    # similar iff k = D*n < k*(n, 0.05)
    exp = {10: 7, 30: 11, 64: 16, 96: 19}
    for n, ks in exp.items():
        tab = np.zeros(n + 1, np.uint8)
        _lib.lib().fringe_ks_table(n, 0.05, tab.ctypes.data)
        assert int(np.argmin(tab)) == ks


def test_ks_table_decides_like_the_oracle_walk():
    # the decision of the integer walk equals the oracle's p >= alpha
    rng = np.random.default_rng(1)
    for n in (10, 30):
        tab = np.zeros(n + 1, np.uint8)
        _lib.lib().fringe_ks_table(n, 0.05, tab.ctypes.data)
        for _ in range(500):
            a = np.sort(rng.rayleigh(size=n).astype(np.float32))
            b = np.sort((rng.rayleigh(size=n) * rng.uniform(0.6, 1.6)).astype(np.float32))
            k = int(round(oracle.ks2_statistic(a, b) * n))
            assert bool(tab[k]) == (oracle.ks2_pvalue(a, b) >= 0.05)


def test_envi_and_vrt_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    rows, cols = 13, 17
    z = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).astype(np.complex64)
    f = str(tmp_path / "a.slc")
    raster.write_envi(f, z)
    assert raster.info(f) == (cols, rows, 1)
    assert np.array_equal(raster.read(f), z)
    assert np.array_equal(raster.read(f, window=(3, 4, 5, 6)), z[4:10, 3:8])
    # raw VRT over the flat file (tops2vrt layout)
    v = str(tmp_path / "a.vrt")
    open(v, "w").write(raster.raw_vrt(f, cols, rows))
    assert np.array_equal(raster.read(v), z)
    # stack VRT with a bbox SrcRect and Date metadata
    s = str(tmp_path / "stack.vrt")
    open(s, "w").write(raster.stack_vrt([(v, "20220101", True), (f, "20220113", False)], cols, rows,
                                        bbox=(2, 10, 1, 9)))
    assert raster.info(s) == (8, 8, 2)
    assert np.array_equal(raster.read(s, 1), z[2:10, 1:9])
    assert np.array_equal(raster.read(s, 2), z[0:8, 0:8])
    assert raster.metadata(s, "Date", "slc", band=2) == "20220113"


def test_vrt_variants_stripmap_and_uavsar(tmp_path):
    """The raw and stack VRTs stripmap2vrt.py (:86-96, :136-152, :163-174) and
    uavsar2vrt.py (:88-98, :124-140, :163-172) write: lower-case
    <sourceFilename> in the per-date raw VRT, a bbox SrcRect in the stack,
    Wavelength/AcquisitionTime metadata, Float64 geometry bands whose raw
    band carries a SrcRect (ignored for raw bands, as GDAL does: the raster
    is the top-left corner), and the UAVSAR LLH file, three float32 per pixel
    read with PixelOffset 12 and ImageOffset 0 / 4 / 8."""
    import os
    rng = np.random.default_rng(5)
    rows, cols = 11, 14
    ymin, ymax, xmin, xmax = 2, 9, 3, 12
    xs, ys = xmax - xmin, ymax - ymin
    bands = []
    for d in ("20200101", "20200113"):
        z = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).astype(np.complex64)
        f = str(tmp_path / (d + ".slc"))
        z.tofile(f)
        with open(str(tmp_path / (d + ".vrt")), "w") as fid:
            fid.write(f'''<VRTDataset rasterXSize="{cols}" rasterYSize="{rows}">
    <VRTRasterBand dataType="CFloat32" band="1" subClass="VRTRawRasterBand">
        <sourceFilename>{os.path.abspath(f)}</sourceFilename>
        <ImageOffset>0</ImageOffset>
        <PixelOffset>8</PixelOffset>
        <LineOffset>{8 * cols}</LineOffset>
        <ByteOrder>LSB</ByteOrder>
    </VRTRasterBand>
</VRTDataset>''')
        bands.append((d, z))
    sv = str(tmp_path / "slcs_base.vrt")
    with open(sv, "w") as fid:
        fid.write(f'<VRTDataset rasterXSize="{xs}" rasterYSize="{ys}">\n')
        for i, (d, _) in enumerate(bands):
            fid.write(f'''    <VRTRasterBand dataType="CFloat32" band="{i + 1}">
        <SimpleSource>
            <SourceFilename>{tmp_path / (d + ".vrt")}</SourceFilename>
            <SourceBand>1</SourceBand>
            <SourceProperties RasterXSize="{cols}" RasterYSize="{rows}" DataType="CFloat32"/>
            <SrcRect xOff="{xmin}" yOff="{ymin}" xSize="{xs}" ySize="{ys}"/>
            <DstRect xOff="0" yOff="0" xSize="{xs}" ySize="{ys}"/>
        </SimpleSource>
        <Metadata domain="slc">
            <MDI key="Date">{d}</MDI>
            <MDI key="Wavelength">0.03</MDI>
            <MDI key="AcquisitionTime">{d}</MDI>
        </Metadata>
    </VRTRasterBand>\n''')
        fid.write("</VRTDataset>")
    assert raster.info(sv) == (xs, ys, 2)
    for i, (d, z) in enumerate(bands):
        assert np.array_equal(raster.read(sv, i + 1), z[ymin:ymax, xmin:xmax])
        assert raster.metadata(sv, "Date", "slc", band=i + 1) == d
        assert raster.metadata(sv, "Wavelength", "slc", band=i + 1) == "0.03"
    # stripmap geometry: Float64 raw band with a SrcRect inside it
    lat = rng.standard_normal((rows, cols))
    lf = str(tmp_path / "lat.rdr")
    lat.tofile(lf)
    gv = str(tmp_path / "lat.vrt")
    with open(gv, "w") as fid:
        fid.write(f'''<VRTDataset rasterXSize="{xs}" rasterYSize="{ys}">
    <VRTRasterBand dataType="Float64" band="1" subClass="VRTRawRasterBand">
        <SourceFilename>{lf}</SourceFilename>
        <ImageOffset>0</ImageOffset>
        <PixelOffset>8</PixelOffset>
        <LineOffset>{8 * cols}</LineOffset>
        <ByteOrder>LSB</ByteOrder>
        <SrcRect xOff="{xmin}" yOff="{ymin}" xSize="{xs}" ySize="{ys}"/>
        <DstRect xOff="0" yOff="0" xSize="{xs}" ySize="{ys}"/>
    </VRTRasterBand>
</VRTDataset>''')
    assert raster.info(gv) == (xs, ys, 1)
    # (read converted to float32, the widest real type the reader hands out)
    assert np.array_equal(raster.read(gv, dtype=np.float32), lat[:ys, :xs].astype(np.float32))
    # UAVSAR LLH: (lat, lon, height) float32 triplets per pixel
    llh = rng.standard_normal((rows, cols, 3)).astype(np.float32)
    hf = str(tmp_path / "x.llh")
    llh.tofile(hf)
    for k in range(3):
        v = str(tmp_path / f"llh{k}.vrt")
        with open(v, "w") as fid:
            fid.write(f'''<VRTDataset rasterXSize="{cols}" rasterYSize="{rows}">
    <VRTRasterBand dataType="Float32" band="1" subClass="VRTRawRasterBand">
        <SourceFilename>{hf}</SourceFilename>
        <ImageOffset>{4 * k}</ImageOffset>
        <PixelOffset>12</PixelOffset>
        <LineOffset>{12 * cols}</LineOffset>
        <ByteOrder>LSB</ByteOrder>
    </VRTRasterBand>
</VRTDataset>''')
        assert np.array_equal(raster.read(v, dtype=np.float32), llh[:, :, k])


def test_multiband_uint32_bip_and_envi_metadata(tmp_path):
    rng = np.random.default_rng(2)
    w = rng.integers(0, 2 ** 32, size=(5, 6, 4), dtype=np.uint64).astype(np.uint32)
    f = str(tmp_path / "wts")
    raster.write_envi(f, w, meta={"HALFWINDOWX": 5, "HALFWINDOWY": 5})
    assert raster.info(f) == (6, 5, 4)
    for b in range(4):
        assert np.array_equal(raster.read(f, b + 1, dtype=np.uint32), w[:, :, b])
    assert raster.metadata(f, "HALFWINDOWX", "ENVI") == "5"
    assert raster.metadata(f, "halfwindowy", "ENVI") == "5"


def test_module_errors_before_gpu_work(tmp_path):
    from fringe_amd.evdlib import Evd
    from fringe_amd.nmaplib import Nmap
    from fringe_amd.phase_linklib import Phaselink
    e = Evd()
    e.verbose = False
    e.inputDS = str(tmp_path / "missing.vrt")
    assert e.run() == 102
    # stack with dates, weights with the wrong window
    rows, cols, n = 6, 7, 4
    slc = synthetic.stack(n, rows, cols, seed=1)
    srcs = []
    for t in range(n):
        d = tmp_path / f"2022010{t + 1}"
        d.mkdir()
        fn = str(d / f"2022010{t + 1}.slc")
        raster.write_envi(fn, slc[t])
        srcs.append((fn, f"2022010{t + 1}", False))
    sv = str(tmp_path / "stack.vrt")
    open(sv, "w").write(raster.stack_vrt(srcs, cols, rows))
    wts = synthetic.full_mask(rows, cols, 2, 2)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": 2, "HALFWINDOWY": 2})
    p = Phaselink()
    p.verbose = False
    p.inputDS, p.weightsDS, p.outputFolder = sv, wf, str(tmp_path / "out")
    p.halfWindowX, p.halfWindowY = 3, 2
    assert p.run() in (108, 109)         # nulong and HALFWINDOWX mismatch
    p.halfWindowX = 2
    p.method, p.bandWidth = "STBAS", 0
    assert p.run() == 101
    p.method = "MLE"
    (tmp_path / "exists").mkdir()
    p.outputFolder = str(tmp_path / "exists")
    assert p.run() == 113
    p.outputFolder = str(tmp_path / "afile")
    open(p.outputFolder, "w").close()
    assert p.run() == 114
    nm = Nmap()
    nm.verbose = False
    nm.inputDS, nm.method = sv, "XX"
    assert nm.run() == 1


def test_server_request_bytes_and_create_without_device():
    L = _lib.lib()
    t = _lib.LinkTile(64, 80, 30, 7, 5, 1, 0, 5, 0, -1, 1, 0, 64)
    r = _lib.ServerRequest(t, 1, None, None, 0.05)
    with_nmap = L.fringe_server_request_bytes(ctypes.byref(r))
    r.d_wts = 1
    given = L.fringe_server_request_bytes(ctypes.byref(r))
    words = L.fringe_mask_words(7, 5)
    npx = 64 * 80
    # the server-computed mask and count are kept with the outputs, and its
    # nmap workspace holds N float amplitudes and a validity byte per pixel,
    # then the forward half of each mask (the centre's word on: position
    # 5 * 15 + 7 = 82 is in word 2, so words 2..5) and a forward count
    fwd_words = words - (5 * 15 + 7) // 32
    extra = (npx * words * 4 + 255) // 256 * 256 + (npx * 4 + 255) // 256 * 256
    extra += ((npx * 30 * 4 + 255) // 256 * 256 + (npx + 255) // 256 * 256
              + (npx * fwd_words * 4 + 255) // 256 * 256 + (npx * 4 + 255) // 256 * 256 + 255) // 256 * 256
    assert with_nmap - given == extra
    outputs = sum((b + 255) // 256 * 256 for b in (npx * 30 * 8, npx * 4, npx * 8))
    assert given >= outputs + L.fringe_link_workspace_bytes(ctypes.byref(t))
    r.prob = 0.0                      # no significance level for the server's nmap
    r.d_wts = None
    assert L.fringe_server_request_bytes(ctypes.byref(r)) == 0
    t.method, t.bandwidth = 2, 0       # STBAS without a bandwidth
    assert L.fringe_server_request_bytes(ctypes.byref(_lib.ServerRequest(t, 1, 1, None, 0.05))) == 0
    h = ctypes.c_void_p()
    rc = L.fringe_server_create(0, 1 << 20, 2, ctypes.byref(h))
    if rc == 0:                        # a GPU is present
        assert L.fringe_server_destroy(h) == 0
    else:
        assert rc == 203 and not h.value
    assert L.fringe_server_create(0, 0, 2, ctypes.byref(h)) == 201


@pytest.mark.parametrize("prob", [0.01, 0.05, 0.1, 0.3])
def test_ad2_threshold_decides_like_the_oracle_pvalue(prob):
    L = _lib.lib()
    sig = ctypes.c_double()
    rng = np.random.default_rng(int(prob * 1000))
    for n in (8, 10, 30, 64, 96):
        thr = L.fringe_ad2_threshold(n, prob, ctypes.byref(sig))
        assert sig.value == oracle.lib().oracle_ad2_sigma(n, n)
        # every statistic within 2000 ulps of the threshold is decided as the
        # p-value decides it (the p-value is monotone there)
        for k in range(-2000, 2001, 7):
            t = float(np.nextafter(thr, np.inf if k > 0 else -np.inf)) if k else thr
            for _ in range(abs(k) - 1 if k else 0):
                t = float(np.nextafter(t, np.inf if k > 0 else -np.inf))
            assert (oracle.ad2_pvalue_of(t) >= prob) == (t <= thr)
        for t in range(200):
            a = rng.gamma(2.0, size=n)
            b = rng.gamma(2.0, size=n) * rng.uniform(0.7, 1.4)
            if t % 3 == 0:
                a, b = np.round(a * 3) / 3, np.round(b * 3) / 3
            a, b = np.sort(a.astype(np.float32)), np.sort(b.astype(np.float32))
            assert (oracle.ad2_statistic(a, b) <= thr) == (oracle.ad2_pvalue(a, b) >= prob)


def test_despeck_module_errors_before_gpu_work(tmp_path):
    from fringe_amd.despecklib import Despeck
    d = Despeck()
    d.verbose = False
    d.inputDS = str(tmp_path / "missing.vrt")
    assert d.run() == 102
    rows, cols, n = 6, 7, 3
    slc = synthetic.stack(n, rows, cols, seed=2)
    srcs = []
    for t in range(n):
        fn = str(tmp_path / f"s{t}.slc")
        raster.write_envi(fn, slc[t])
        srcs.append((fn, f"2022010{t + 1}", False))
    sv = str(tmp_path / "stack.vrt")
    open(sv, "w").write(raster.stack_vrt(srcs, cols, rows))
    d.inputDS = sv
    d.band1 = 4
    assert d.run() == 102                    # band outside the stack
    d.band1, d.band2 = 1, 5
    assert d.run() == 102
    d.band2 = 2
    d.weightsDS = ""                         # the reference needs the mask
    assert d.run() == 105
    wts = synthetic.full_mask(rows, cols, 2, 2)
    wf = str(tmp_path / "wts")
    raster.write_envi(wf, wts, meta={"HALFWINDOWX": 2, "HALFWINDOWY": 2})
    d.weightsDS = wf
    d.halfWindowX, d.halfWindowY = 3, 2
    assert d.run() == 109                    # last failing check wins, as in despeck.cpp:90-140
    d.halfWindowX, d.halfWindowY = 2, 1
    assert d.run() == 110


def test_calamp_module_errors_before_gpu_work(tmp_path):
    from fringe_amd.calamplib import Calamp
    c = Calamp()
    c.verbose = False
    c.inputDS = str(tmp_path / "missing.vrt")
    c.outputDS = str(tmp_path / "out.vrt")
    assert c.run() == 102
    rows, cols = 5, 6
    slc = synthetic.stack(2, rows, cols, seed=4)
    srcs = []
    for t in range(2):
        fn = str(tmp_path / f"s{t}.slc")
        raster.write_envi(fn, slc[t])
        srcs.append((fn, f"2022010{t + 1}", False))
    sv = str(tmp_path / "stack.vrt")
    open(sv, "w").write(raster.stack_vrt(srcs, cols, rows))
    c.inputDS = sv
    c.maskDS = str(tmp_path / "nomask")
    assert c.run() == 103
    mf = str(tmp_path / "mask")
    raster.write_envi(mf, np.ones((rows, cols + 1), np.uint8))
    c.maskDS = mf
    assert c.run() == 104


def test_schedule_table_loads():
    """The committed autotuner table (fringe_amd/schedules.json) names valid
    keys and values, and loading it installs them in the library's table."""
    import json
    from fringe_amd import _lib
    L = _lib.lib()
    with open(_lib.SCHEDULE_FILE) as f:
        t = json.load(f)
    assert t["gram"] and t["solve"]
    for e in t["gram"]:
        assert e["variant"] in _lib.GRAM_VARIANTS and e["variant"] == min(e["ms"], key=e["ms"].get)
    for e in t["solve"]:
        assert e["launch"] in ("grid", "persistent") and e["mode"] in (0, 1, 2) and e["precision"] in (0, 1)
    L.fringe_schedule_clear()
    assert _lib.load_schedules(L) == len(t["gram"]) + len(t["solve"]) + len(t.get("generic_a", []))
    for e in t.get("generic_a", []):
        assert L.fringe_schedule_get(_lib.FRINGE_SCHED_GENERIC_A, e["n"], e["precision"], 0) == \
            _lib.GENERIC_A_PLACES[e["work_matrix"]]
    for e in t["gram"]:
        assert L.fringe_schedule_get(_lib.FRINGE_SCHED_GRAM_VARIANT, e["n"], e["Nx"], e["Ny"]) == \
            _lib.GRAM_VARIANTS[e["variant"]]
    for e in t["solve"]:
        want = 0 if e["launch"] == "grid" else (e["waves_per_cu"] << 8) | e["grab"]
        assert L.fringe_schedule_get(_lib.FRINGE_SCHED_SOLVE, e["n"], e["mode"], e["precision"]) == want
    assert L.fringe_schedule_set(_lib.FRINGE_SCHED_GRAM_VARIANT, 30, 7, 5, 9) == 201
    assert L.fringe_schedule_set(7, 30, 7, 5, 0) == 201
    # rows in registers: 33..128 dates, either precision; no work matrix in
    # global memory up to 64 dates
    G = _lib.FRINGE_SCHED_GENERIC_A
    assert L.fringe_schedule_set(G, 160, 1, 0, 2) == 201 and L.fringe_schedule_set(G, 160, 0, 0, 2) == 201
    assert L.fringe_schedule_set(G, 96, 1, 0, 3) == 201 and L.fringe_schedule_set(G, 96, 0, 0, 2) == 0
    assert L.fringe_schedule_set(G, 64, 1, 0, 1) == 201 and L.fringe_schedule_set(G, 48, 0, 0, 2) == 0
    L.fringe_schedule_clear()
    assert L.fringe_schedule_get(G, 80, 1, 0) == 2 and L.fringe_schedule_get(G, 80, 0, 0) == 2
    assert L.fringe_schedule_get(G, 128, 1, 0) == 2 and L.fringe_schedule_get(G, 160, 1, 0) == 0
    # defaults up to 64 dates: rows for FP64 above 32 dates and for FP32 at
    # 33..48; one row per lane for FP32 at 49..64 and for both up to 32
    assert [L.fringe_schedule_get(G, n, 0, 0) for n in (30, 33, 48, 49, 64)] == [0, 2, 2, 2, 2]
    assert [L.fringe_schedule_get(G, n, 1, 0) for n in (30, 33, 48, 49, 64)] == [0, 2, 2, 0, 0]
    assert L.fringe_schedule_set(G, 72, 1, 0, 0) == 0 and L.fringe_schedule_get(G, 96, 1, 0) == 0
    assert L.fringe_schedule_get(G, 64, 1, 0) == 0 and L.fringe_schedule_get(G, 48, 1, 0) == 2
    assert L.fringe_schedule_set(G, 40, 0, 0, 0) == 0 and L.fringe_schedule_get(G, 48, 0, 0) == 0
    assert L.fringe_schedule_get(G, 64, 0, 0) == 2
    L.fringe_schedule_clear()
    _lib.load_schedules(L)

