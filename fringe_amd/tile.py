"""Tile engine on device-resident buffers (torch tensors used as HBM storage).

``link`` and ``nmap`` call the C ABI of libfringe_amd.so on a HIP stream;
torch only allocates device memory and provides the stream.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_DTYPE_U32 = torch.int32   # bitmask words are stored as int32 bit patterns


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}: {_lib.last_error()}")


def mask_words(Nx, Ny):
    return _lib.lib().fringe_mask_words(Nx, Ny)


def ks_table(n, prob):
    """(table[k] for k = 0..n, smallest |p_k - prob|)."""
    tab = np.zeros(n + 1, np.uint8)
    margin = _lib.lib().fringe_ks_table(n, prob, tab.ctypes.data)
    return tab, margin


def link_params(rows, cols, nbands, Nx, Ny, module="phase_link", method="MLE",
                min_neighbors=2, ref_index=0, bandwidth=-1, precision="mixed",
                row_begin=0, row_end=None):
    t = _lib.LinkTile()
    t.rows, t.cols, t.nbands, t.Nx, t.Ny = rows, cols, nbands, Nx, Ny
    t.module = _lib.FRINGE_MODULE_EVD if module == "evd" else _lib.FRINGE_MODULE_PHASE_LINK
    t.method = _lib.METHODS[method]
    t.min_neighbors = min_neighbors
    t.ref_index = ref_index
    t.bandwidth = bandwidth
    t.precision = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    t.row_begin = row_begin
    t.row_end = rows if row_end is None else row_end
    return t


class LinkRunner:
    """Preallocated outputs + workspace for repeated phase linking of one
    tile shape (the request loop and the benchmark reuse it)."""

    def __init__(self, params, device="cuda"):
        self.p = params
        L = _lib.lib()
        self.work_bytes = L.fringe_link_workspace_bytes(ctypes.byref(params))
        if self.work_bytes == 0:
            raise RuntimeError("invalid or unsupported link tile parameters")
        dev = torch.device(device)
        n, r, c = params.nbands, params.rows, params.cols
        self.work = torch.empty(self.work_bytes, dtype=torch.uint8, device=dev)
        self.phase = torch.zeros((n, r, c), dtype=torch.complex64, device=dev)
        self.tcorr = torch.zeros((r, c), dtype=torch.float32, device=dev)
        self.comp = torch.zeros((r, c), dtype=torch.complex64, device=dev)

    def run(self, slc, wts, stream=None):
        assert slc.is_contiguous() and slc.dtype == torch.complex64
        assert wts.is_contiguous() and wts.dtype == _DTYPE_U32
        p = self.p
        assert tuple(slc.shape) == (p.nbands, p.rows, p.cols)
        assert wts.numel() == p.rows * p.cols * mask_words(p.Nx, p.Ny)
        rc = _lib.lib().fringe_link_run(ctypes.byref(p), _ptr(slc), _ptr(wts), _ptr(self.phase),
                                        _ptr(self.tcorr), _ptr(self.comp), _ptr(self.work),
                                        self.work_bytes, _stream(stream))
        _check(rc, "fringe_link_run")
        return self.phase, self.tcorr, self.comp

    def resolved_pixels(self):
        """Linear indices (row * cols + col) of the pixels that the last run
        recomputed with the FP64 path because their FP32 result was not
        certified (mixed precisions; empty otherwise).  Synchronises."""
        torch.cuda.synchronize(self.work.device)
        cnt = _lib.lib().fringe_link_resolved_pixels(ctypes.byref(self.p), _ptr(self.work), self.work_bytes)
        if cnt < 0:
            raise RuntimeError("fringe_link_resolved_pixels failed")
        if cnt == 0:
            return np.zeros(0, np.int64)
        lst = self.work[256:256 + 4 * cnt].view(torch.int32).cpu().numpy()
        return np.sort(lst.astype(np.int64))


class BatchRunner:
    """Outputs + workspace for fringe_link_batch_run over a list of
    (slc, wts, params) tiles of different date counts and windows (one
    module, method and precision); the solves of each date class run in one
    persistent launch per round."""

    def __init__(self, tiles, device="cuda"):
        L = _lib.lib()
        dev = torch.device(device)
        self.tiles = tiles
        self.outs = []
        arr = (_lib.BatchItem * len(tiles))()
        for i, (slc, wts, p) in enumerate(tiles):
            assert slc.is_contiguous() and slc.dtype == torch.complex64
            assert wts.is_contiguous() and wts.dtype == _DTYPE_U32
            assert tuple(slc.shape) == (p.nbands, p.rows, p.cols)
            assert wts.numel() == p.rows * p.cols * mask_words(p.Nx, p.Ny)
            ph = torch.zeros((p.nbands, p.rows, p.cols), dtype=torch.complex64, device=dev)
            tc = torch.zeros((p.rows, p.cols), dtype=torch.float32, device=dev)
            co = torch.zeros((p.rows, p.cols), dtype=torch.complex64, device=dev)
            self.outs.append((ph, tc, co))
            arr[i].tile = p
            arr[i].d_slc, arr[i].d_wts = slc.data_ptr(), wts.data_ptr()
            arr[i].d_phase, arr[i].d_tcorr, arr[i].d_comp = ph.data_ptr(), tc.data_ptr(), co.data_ptr()
        self.items = arr
        self.work_bytes = L.fringe_link_batch_workspace_bytes(arr, len(tiles))
        if self.work_bytes == 0:
            raise RuntimeError("invalid or unsupported batch (one module, method and precision per batch)")
        self.work = torch.empty(self.work_bytes, dtype=torch.uint8, device=dev)

    def run(self, stream=None):
        rc = _lib.lib().fringe_link_batch_run(self.items, len(self.tiles), _ptr(self.work), self.work_bytes,
                                              _stream(stream))
        _check(rc, "fringe_link_batch_run")
        return self.outs


def link(slc, wts, Nx, Ny, **kw):
    """Phase-link a device tile.  slc: complex64 [n, rows, cols] on the GPU;
    wts: int32 [rows, cols, nulong] on the GPU.  Returns (phase, tcorr, comp)."""
    n, rows, cols = slc.shape
    runner = LinkRunner(link_params(rows, cols, n, Nx, Ny, **kw), device=slc.device)
    return runner.run(slc, wts)


def coherence(slc, wts, Nx, Ny, precision="mixed", module="phase_link", min_neighbors=2, row_begin=0,
              row_end=None, stream=None):
    """Coherence matrices of a device tile's output rows: (gamma [pixels, n, n]
    complex64 (MFMA Gram) or complex128, npix [pixels] int32, -1 = not
    linked)."""
    n, rows, cols = slc.shape
    p = link_params(rows, cols, n, Nx, Ny, module=module, min_neighbors=min_neighbors, precision=precision,
                    row_begin=row_begin, row_end=row_end)
    L = _lib.lib()
    wb = L.fringe_coherence_workspace_bytes(ctypes.byref(p))
    if wb == 0:
        raise RuntimeError("invalid or unsupported coherence parameters")
    work = torch.empty(wb, dtype=torch.uint8, device=slc.device)
    npx = (p.row_end - p.row_begin) * cols
    c64 = p.precision != _lib.FRINGE_PREC_FP64 and n <= 128
    gam = torch.empty((npx, n, n), dtype=torch.complex64 if c64 else torch.complex128, device=slc.device)
    npix = torch.empty(npx, dtype=torch.int32, device=slc.device)
    rc = L.fringe_coherence_run(ctypes.byref(p), _ptr(slc.contiguous()), _ptr(wts.contiguous()), _ptr(gam),
                                _ptr(npix), _ptr(work), wb, _stream(stream))
    _check(rc, "fringe_coherence_run")
    return gam, npix


def nmap(slc, Nx, Ny, prob=0.05, mask=None, amp_const=None, row_begin=0, row_end=None,
         stream=None, method="KS2"):
    """SHP map of a device tile (method "KS2" or "AD2").  Returns
    (wts int32 [rows, cols, nulong], count int32 [rows, cols])."""
    n, rows, cols = slc.shape
    t = _lib.NmapTile()
    t.rows, t.cols, t.nbands, t.Nx, t.Ny = rows, cols, n, Nx, Ny
    t.prob, t.method = prob, {"KS2": 0, "AD2": 1}[method]
    t.row_begin, t.row_end = row_begin, rows if row_end is None else row_end
    L = _lib.lib()
    wb = L.fringe_nmap_workspace_bytes(ctypes.byref(t))
    work = torch.empty(max(wb, 1), dtype=torch.uint8, device=slc.device)
    nul = mask_words(Nx, Ny)
    wts = torch.zeros((rows, cols, nul), dtype=_DTYPE_U32, device=slc.device)
    cnt = torch.zeros((rows, cols), dtype=torch.int32, device=slc.device)
    ac = None
    if amp_const is not None:
        ac = np.ascontiguousarray(amp_const, np.float64)
    m = None
    if mask is not None:
        m = mask.to(device=slc.device, dtype=torch.uint8).contiguous()
    rc = L.fringe_nmap_run(ctypes.byref(t), _ptr(slc.contiguous()), _ptr(m),
                           None if ac is None else ctypes.c_void_p(ac.ctypes.data),
                           _ptr(wts), _ptr(cnt), _ptr(work), wb, _stream(stream))
    _check(rc, "fringe_nmap_run")
    return wts, cnt
