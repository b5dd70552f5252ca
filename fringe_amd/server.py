"""Tile server: continuous batching of phase-linking requests on one GPU.

Wraps the fringe_server_* C ABI.  Each request names a device-resident stack
(a torch tensor used as HBM storage) and, optionally, its SHP mask; the
server admits requests oldest first as streams and arena space free up, runs
admitted requests concurrently on its own streams, and keeps their outputs in
its arena until they are fetched and released.
"""
import ctypes

import torch

from . import _lib
from .tile import link_params, mask_words

_HIP = None


def _hip():
    """The HIP runtime already loaded in this process (shared with torch)."""
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p]
    return _HIP


_D2D = 3   # hipMemcpyDeviceToDevice


class TileServer:
    def __init__(self, budget_bytes, streams=4, device=0):
        self.L = _lib.lib()
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        rc = self.L.fringe_server_create(device, int(budget_bytes), streams, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"fringe_server_create failed with code {rc}: {_lib.last_error()}")
        self.h = h
        self._keep = {}    # id -> (tensors the request reads, shape info)

    def request_bytes(self, params, with_mask=True):
        r = _lib.ServerRequest()
        r.tile = params
        r.d_slc = ctypes.c_void_p(1)
        r.d_wts = ctypes.c_void_p(1) if with_mask else None
        r.prob = 0.05
        return self.L.fringe_server_request_bytes(ctypes.byref(r))

    def submit(self, slc, params, wts=None, prob=0.05, mask=None):
        """Queue one tile; returns the request id.  `slc` ([n][rows][cols]
        complex64) and `wts` must stay unchanged until the request is done."""
        if slc.device != self.device or slc.dtype != torch.complex64 or not slc.is_contiguous():
            raise ValueError("slc must be a contiguous complex64 tensor on the server's device")
        n, rows, cols = slc.shape
        if (params.nbands, params.rows, params.cols) != (n, rows, cols):
            raise ValueError("params do not match the stack shape")
        if wts is not None and (wts.shape != (rows, cols, mask_words(params.Nx, params.Ny))
                                or not wts.is_contiguous()):
            raise ValueError("wts must be [rows][cols][words] and contiguous")
        r = _lib.ServerRequest()
        r.tile = params
        r.d_slc = ctypes.c_void_p(slc.data_ptr())
        r.d_wts = ctypes.c_void_p(wts.data_ptr()) if wts is not None else None
        r.d_mask = ctypes.c_void_p(mask.data_ptr()) if mask is not None else None
        r.prob = prob
        # the request is read on the server's streams: finish the producer's work first
        torch.cuda.current_stream(self.device).synchronize()
        rid = ctypes.c_longlong()
        rc = self.L.fringe_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid))
        if rc:
            raise RuntimeError(f"fringe_server_submit failed with code {rc}")
        self._keep[rid.value] = ((slc, wts, mask), (n, rows, cols, params.Nx, params.Ny))
        return rid.value

    def poll(self, rid):
        st = ctypes.c_int()
        code = self.L.fringe_server_poll(self.h, rid, ctypes.byref(st))
        return st.value, code

    def wait(self, rid):
        return self.L.fringe_server_wait(self.h, rid)

    def times(self, rid):
        q, t = ctypes.c_double(), ctypes.c_double()
        rc = self.L.fringe_server_times(self.h, rid, ctypes.byref(q), ctypes.byref(t))
        if rc:
            raise RuntimeError(f"request {rid} is not done")
        return q.value, t.value

    def fetch(self, rid):
        """Copy a done request's outputs into new tensors:
        (phase, tcorr, comp, wts or None, count or None).  Returns after the
        copies have finished, so the request may be released at once."""
        ptrs = [ctypes.c_void_p() for _ in range(5)]
        code = self.L.fringe_server_result(self.h, rid, *[ctypes.byref(p) for p in ptrs])
        if code:
            raise RuntimeError(f"request {rid} failed with code {code}")
        (_, (n, rows, cols, Nx, Ny)) = self._keep[rid]
        dev = self.device
        out = [torch.empty((n, rows, cols), dtype=torch.complex64, device=dev),
               torch.empty((rows, cols), dtype=torch.float32, device=dev),
               torch.empty((rows, cols), dtype=torch.complex64, device=dev)]
        if ptrs[3].value:
            out.append(torch.empty((rows, cols, mask_words(Nx, Ny)), dtype=torch.int32, device=dev))
            out.append(torch.empty((rows, cols), dtype=torch.int32, device=dev))
        else:
            out += [None, None]
        s = torch.cuda.current_stream(dev)
        H = _hip()
        for p, t in zip(ptrs, out):
            if t is not None:
                e = H.hipMemcpyAsync(ctypes.c_void_p(t.data_ptr()), p, t.numel() * t.element_size(), _D2D,
                                     ctypes.c_void_p(s.cuda_stream))
                if e:
                    raise RuntimeError(f"hipMemcpyAsync failed ({e})")
        # the copies must have read the arena before release() can hand the
        # region to another request, whose kernels run on a server stream
        # that is not ordered with this one
        s.synchronize()
        return tuple(out)

    def release(self, rid):
        rc = self.L.fringe_server_release(self.h, rid)
        if rc:
            raise RuntimeError(f"release of request {rid} failed ({rc})")
        self._keep.pop(rid, None)

    def close(self):
        if self.h:
            rc = self.L.fringe_server_destroy(self.h)
            self.h = None
            self._keep.clear()
            if rc:
                raise RuntimeError(f"fringe_server_destroy failed ({rc})")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["TileServer", "link_params"]
