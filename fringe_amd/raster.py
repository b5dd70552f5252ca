"""Raster helpers for the drivers and tests: ENVI writing in numpy, VRT text
writers (per-file raw VRTs and stack VRTs in the layout tops2vrt.py and
Stack.py write), and reads through libfringe_amd's ENVI/VRT reader."""
import ctypes
import os

import numpy as np

from . import _lib

ENVI_CODES = {np.dtype(np.uint8): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 3,
              np.dtype(np.float32): 4, np.dtype(np.float64): 5, np.dtype(np.complex64): 6,
              np.dtype(np.complex128): 9, np.dtype(np.uint16): 12, np.dtype(np.uint32): 13}
GDAL_NAMES = {np.dtype(np.uint8): "Byte", np.dtype(np.int16): "Int16", np.dtype(np.int32): "Int32",
              np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
              np.dtype(np.complex64): "CFloat32", np.dtype(np.uint32): "UInt32"}


def write_envi(path, data, interleave="bip", meta=None):
    """Write [rows, cols] or [rows, cols, bands] (bip) as ENVI with <path>.hdr."""
    a = np.ascontiguousarray(data)
    bands = 1 if a.ndim == 2 else a.shape[2]
    rows, cols = a.shape[:2]
    a.tofile(path)
    with open(path + ".hdr", "w") as f:
        f.write("ENVI\ndescription = {\n%s}\n" % path)
        f.write(f"samples = {cols}\nlines   = {rows}\nbands   = {bands}\nheader offset = 0\n")
        f.write(f"file type = ENVI Standard\ndata type = {ENVI_CODES[a.dtype]}\n")
        f.write(f"interleave = {interleave}\nbyte order = 0\n")
        for k, v in (meta or {}).items():
            f.write(f"{k} = {v}\n")


def raw_vrt(path, cols, rows, dtype=np.complex64):
    """Single-band VRTRawRasterBand over a flat binary file (tops2vrt.py:144-180)."""
    es = np.dtype(dtype).itemsize
    return (f'<VRTDataset rasterXSize="{cols}" rasterYSize="{rows}">\n'
            f'    <VRTRasterBand dataType="{GDAL_NAMES[np.dtype(dtype)]}" band="1" subClass="VRTRawRasterBand">\n'
            f'        <SourceFilename>{os.path.abspath(path)}</SourceFilename>\n'
            f'        <ImageOffset>0</ImageOffset>\n        <PixelOffset>{es}</PixelOffset>\n'
            f'        <LineOffset>{es * cols}</LineOffset>\n        <ByteOrder>LSB</ByteOrder>\n'
            f'    </VRTRasterBand>\n</VRTDataset>')


def stack_vrt(sources, width, height, bbox=None, metadata=None):
    """Stack VRT of SimpleSource bands with Date metadata in domain "slc"
    (tops2vrt.py:196-225, Stack.py:77-157).  sources: list of
    (vrt_or_raster_path, date, apply_bbox); bbox = (ymin, ymax, xmin, xmax)."""
    ymin, ymax, xmin, xmax = bbox if bbox else (0, height, 0, width)
    xs, ys = xmax - xmin, ymax - ymin
    out = [f'<VRTDataset rasterXSize="{xs}" rasterYSize="{ys}">']
    for i, (path, date, use_bbox) in enumerate(sources):
        x0, y0 = (xmin, ymin) if (bbox and use_bbox) else (0, 0)
        md = "".join(f'\n            <MDI key="{k}">{v}</MDI>' for k, v in (metadata[i] if metadata else {}).items())
        out.append(f'''    <VRTRasterBand dataType="CFloat32" band="{i + 1}">
        <SimpleSource>
            <SourceFilename>{os.path.abspath(path)}</SourceFilename>
            <SourceBand>1</SourceBand>
            <SourceProperties RasterXSize="{width}" RasterYSize="{height}" DataType="CFloat32"/>
            <SrcRect xOff="{x0}" yOff="{y0}" xSize="{xs}" ySize="{ys}"/>
            <DstRect xOff="0" yOff="0" xSize="{xs}" ySize="{ys}"/>
        </SimpleSource>
        <Metadata domain="slc">
            <MDI key="Date">{date}</MDI>{md}
        </Metadata>
    </VRTRasterBand>''')
    out.append("</VRTDataset>")
    return "\n".join(out)


def envi_vrt(path, cols, rows, bands, dtype):
    """VRT describing one BIP ENVI file band by band (what evd.py writes with
    gdal.Translate for every output, evd.py:69-76)."""
    es = np.dtype(dtype).itemsize
    out = [f'<VRTDataset rasterXSize="{cols}" rasterYSize="{rows}">']
    for b in range(bands):
        out.append(f'  <VRTRasterBand dataType="{GDAL_NAMES[np.dtype(dtype)]}" band="{b + 1}" subClass="VRTRawRasterBand">\n'
                   f'    <SourceFilename relativeToVRT="0">{os.path.abspath(path)}</SourceFilename>\n'
                   f'    <ImageOffset>{es * b}</ImageOffset>\n    <PixelOffset>{es * bands}</PixelOffset>\n'
                   f'    <LineOffset>{es * bands * cols}</LineOffset>\n    <ByteOrder>LSB</ByteOrder>\n  </VRTRasterBand>')
    out.append("</VRTDataset>")
    return "\n".join(out)


def info(path):
    """(cols, rows, bands) of an ENVI or VRT raster."""
    x, y, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = _lib.lib().fringe_raster_info(path.encode(), ctypes.byref(x), ctypes.byref(y), ctypes.byref(b))
    if rc != 0:
        raise OSError(f"cannot open raster {path}")
    return x.value, y.value, b.value


def metadata(path, key, domain, band=0):
    buf = ctypes.create_string_buffer(512)
    rc = _lib.lib().fringe_raster_metadata(path.encode(), band, key.encode(), domain.encode(), buf, 512)
    if rc != 0:
        raise OSError(f"cannot open raster {path}")
    return buf.value.decode()


_TYPES = {np.complex64: 0, np.float32: 1, np.uint32: 2, np.uint8: 3}


def read(path, band=1, window=None, dtype=np.complex64):
    """Read one band (1-based) as a [rows, cols] array of dtype."""
    cols, rows, _ = info(path)
    x0, y0, w, h = window if window else (0, 0, cols, rows)
    out = np.zeros((h, w), dtype)
    rc = _lib.lib().fringe_raster_read(path.encode(), band, x0, y0, w, h, _TYPES[dtype], out.ctypes.data)
    if rc != 0:
        raise OSError(f"cannot read {path} band {band}")
    return out
