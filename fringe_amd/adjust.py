"""Datum adjustment of sequential mini-stacks (python/adjustMiniStacks.py).

The reference writes, for every date, a VRT whose pixel function combines
the date's mini-stack phase with the datum-connection phase of its
mini-stack: "mul" of the wrapped phases (adjustMiniStacks.py:180-203, :206-228)
or "sum" of the unwrapped ones (:157-178, with a time-series VRT whose
ScaleOffset removes the phase at a reference pixel, :231-270).  Here the same
combination is computed on the GPU (fringe_adjust_wrapped /
fringe_adjust_unwrapped, evaluated in double and rounded to the band type as
GDAL evaluates the pixel functions) and written as ENVI rasters with a VRT
beside each, under the names the reference gives its VRTs.
"""
import datetime
import glob
import os
import time

import numpy as np

from . import _lib, raster

_BLOCK_ROWS = 2048


def get_dates(slcDir):
    """Sorted dates of the <date>.vrt files in slcDir (adjustMiniStacks.py:54-64)."""
    return sorted(os.path.basename(p).replace(".vrt", "") for p in glob.glob(os.path.join(slcDir, "*.vrt")))


def get_stack_dict(dateList, miniStackDir, datumDir, outDir, miniStackSize, subDir="EVD",
                   fileExtension=".slc", outputExtension=".phase", mini_stack_sizes=None):
    """date -> [mini-stack phase, datum phase, temporal coherence, output]
    (adjustMiniStacks.py:66-105).  mini_stack_sizes (optional) gives uneven
    mini-stack lengths, as the sequential driver accepts."""
    miniStackDir, datumDir, outDir = (os.path.abspath(p) for p in (miniStackDir, datumDir, outDir))
    stackDict = {}
    indStart, k = 0, 0
    while indStart < len(dateList):
        size = miniStackSize if mini_stack_sizes is None else mini_stack_sizes[k % len(mini_stack_sizes)]
        k += 1
        indEnd = min(indStart + size, len(dateList))
        path = os.path.join(miniStackDir, dateList[indStart] + "_" + dateList[indEnd - 1])
        datum = os.path.join(datumDir, subDir, dateList[indEnd - 1] + fileExtension)
        for dd in dateList[indStart:indEnd]:
            stackDict[dd] = [os.path.join(path, subDir, dd + fileExtension), datum,
                             os.path.join(path, "EVD", "tcorr.bin"), os.path.join(outDir, dd + outputExtension)]
        indStart = indEnd
    return stackDict


def _readable(path):
    """The reference reads <file>.vrt; fall back to the ENVI file itself."""
    return path + ".vrt" if os.path.exists(path + ".vrt") else path


def _combine(a_path, b_path, out_path, dtype, entry):
    import torch
    cols, rows, _ = raster.info(a_path)
    cb, rb, _ = raster.info(b_path)
    if (cols, rows) != (cb, rb):
        raise ValueError(f"{a_path} and {b_path} differ in size")
    L = _lib.lib()
    dev = torch.device("cuda")
    tdt = torch.complex64 if dtype == np.complex64 else torch.float32
    out = np.zeros((rows, cols), dtype)
    stream = torch.cuda.current_stream()
    for y0 in range(0, rows, _BLOCK_ROWS):
        h = min(_BLOCK_ROWS, rows - y0)
        a = torch.from_numpy(raster.read(a_path, window=(0, y0, cols, h), dtype=dtype)).to(dev)
        b = torch.from_numpy(raster.read(b_path, window=(0, y0, cols, h), dtype=dtype)).to(dev)
        o = torch.empty((h, cols), dtype=tdt, device=dev)
        rc = getattr(L, entry)(a.data_ptr(), b.data_ptr(), o.data_ptr(), h * cols, stream.cuda_stream)
        if rc:
            raise RuntimeError(f"{entry} failed with code {rc}: {_lib.last_error()}")
        out[y0:y0 + h] = o.cpu().numpy()
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    raster.write_envi(out_path, out)
    with open(out_path + ".vrt", "w") as fid:
        fid.write(raster.envi_vrt(out_path, cols, rows, 1, dtype))
    return out


def adjust_wrapped(dateList, miniStackDir, datumDir, outDir, miniStackSize, mini_stack_sizes=None,
                   verbose=True):
    """<outDir>/<date>.slc = mini-stack phase x datum phase (CFloat32)."""
    sd = get_stack_dict(dateList, miniStackDir, datumDir, outDir, miniStackSize, subDir="EVD",
                        fileExtension=".slc", outputExtension=".slc", mini_stack_sizes=mini_stack_sizes)
    for k, (mini, datum, _, output) in sd.items():
        if verbose:
            print(f"{k}: {mini} x {datum} -> {output}")
        _combine(_readable(mini), _readable(datum), output, np.complex64, "fringe_adjust_wrapped")
    return sd


def _acquisition_time(date):
    d1 = datetime.datetime(*time.strptime(date, "%Y%m%d")[0:6])
    return float(d1.year) + float(d1.timetuple().tm_yday - 1) / 365.25


def adjust_unwrapped(dateList, miniStackDir, datumDir, outDir, miniStackSize, timeSeriesName, ref_x, ref_y,
                     mini_stack_sizes=None, verbose=True):
    """<outDir>/<date>.phase = mini-stack + datum unwrapped phase (Float32),
    and a time-series VRT whose bands carry ScaleOffset = minus the phase at
    (ref_y, ref_x) (adjustMiniStacks.py:231-270)."""
    sd = get_stack_dict(dateList, miniStackDir, datumDir, outDir, miniStackSize, subDir="unwrap",
                        fileExtension=".unw", outputExtension=".phase", mini_stack_sizes=mini_stack_sizes)
    bands = []
    width = length = 0
    for k, (mini, datum, _, output) in sd.items():
        if verbose:
            print(f"{k}: {mini} + {datum} -> {output}")
        out = _combine(_readable(mini), _readable(datum), output, np.float32, "fringe_adjust_unwrapped")
        length, width = out.shape
        ref = -1 * out[ref_y, ref_x]
        bands.append(f'''<VRTRasterBand dataType="Float32">
        <SimpleSource>
            <SourceFilename>{output}.vrt</SourceFilename>
        </SimpleSource>
        <Metadata domain="slc">
            <MDI key="Date">{k}</MDI>
            <MDI key="AcquisitionTime">{_acquisition_time(k)}</MDI>
        </Metadata>
        <ScaleOffset>{ref}</ScaleOffset>
</VRTRasterBand>\n''')
    if timeSeriesName:
        with open(timeSeriesName, "w") as fid:
            fid.write(f'<VRTDataset rasterXSize="{width}" rasterYSize="{length}">\n')
            fid.write("".join(bands))
            fid.write("</VRTDataset>")
    return sd
