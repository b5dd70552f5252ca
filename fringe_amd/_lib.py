"""ctypes binding of libfringe_amd.so (the C ABI in include/fringe_amd.h).

The library is built in-tree (``python -m fringe_amd.build``).  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfringe_amd.so")

FRINGE_MODULE_EVD, FRINGE_MODULE_PHASE_LINK = 0, 1
FRINGE_METHOD_MLE, FRINGE_METHOD_EVD, FRINGE_METHOD_STBAS = 0, 1, 2
FRINGE_PREC_FP64, FRINGE_PREC_MIXED, FRINGE_PREC_MIXED_BF16 = 0, 1, 2
METHODS = {"MLE": FRINGE_METHOD_MLE, "EVD": FRINGE_METHOD_EVD, "STBAS": FRINGE_METHOD_STBAS}
PRECISIONS = {"fp64": FRINGE_PREC_FP64, "mixed": FRINGE_PREC_MIXED, "mixed_bf16": FRINGE_PREC_MIXED_BF16}

# Symbols declared in include/fringe_amd.h
EXPORTS = [
    "fringe_evd_options_init", "fringe_evd_process", "fringe_phase_link_process",
    "fringe_nmap_options_init", "fringe_nmap_process",
    "fringe_link_workspace_bytes", "fringe_link_run",
    "fringe_nmap_workspace_bytes", "fringe_nmap_run", "fringe_ks_table", "fringe_ad2_threshold",
    "fringe_mask_words", "fringe_version", "fringe_last_error",
    "fringe_timing_enable", "fringe_timing_reset", "fringe_timing_read", "fringe_timing_stage_name",
    "fringe_raster_info", "fringe_raster_metadata", "fringe_raster_read",
    "fringe_server_request_bytes", "fringe_server_create", "fringe_server_submit", "fringe_server_poll",
    "fringe_server_wait", "fringe_server_result", "fringe_server_times", "fringe_server_release",
    "fringe_server_destroy", "fringe_adjust_wrapped", "fringe_adjust_unwrapped",
    "fringe_despeck_options_init", "fringe_despeck_process",
    "fringe_calamp_options_init", "fringe_calamp_process",
    "fringe_coherence_workspace_bytes", "fringe_coherence_run", "fringe_link_resolved_pixels",
    "fringe_link_run_noalloc", "fringe_link_batch_workspace_bytes", "fringe_link_batch_run",
    "fringe_schedule_set", "fringe_schedule_get", "fringe_schedule_clear",
]
FRINGE_REQ_QUEUED, FRINGE_REQ_RUNNING, FRINGE_REQ_DONE = 0, 1, 2


class LinkTile(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in (
        "rows", "cols", "nbands", "Nx", "Ny", "module", "method", "min_neighbors",
        "ref_index", "bandwidth", "precision", "row_begin", "row_end")]


class NmapTile(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int), ("cols", ctypes.c_int), ("nbands", ctypes.c_int),
                ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("prob", ctypes.c_double),
                ("method", ctypes.c_int), ("row_begin", ctypes.c_int), ("row_end", ctypes.c_int)]


class EvdOptions(ctypes.Structure):
    _fields_ = [("inputDS", ctypes.c_char_p), ("wtsDS", ctypes.c_char_p),
                ("compSlc", ctypes.c_char_p), ("outputCompressedSlcFolder", ctypes.c_char_p),
                ("outputFolder", ctypes.c_char_p), ("blocksize", ctypes.c_int),
                ("memsize", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int),
                ("minNeighbors", ctypes.c_int), ("miniStackCount", ctypes.c_int),
                ("bandWidth", ctypes.c_int), ("method", ctypes.c_char_p),
                ("precision", ctypes.c_int), ("verbose", ctypes.c_int)]


class NmapOptions(ctypes.Structure):
    _fields_ = [("inputDS", ctypes.c_char_p), ("maskDS", ctypes.c_char_p),
                ("wtsDS", ctypes.c_char_p), ("ncountDS", ctypes.c_char_p),
                ("method", ctypes.c_char_p), ("noGPU", ctypes.c_int), ("blocksize", ctypes.c_int),
                ("memsize", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int),
                ("prob", ctypes.c_double), ("verbose", ctypes.c_int)]


class DespeckOptions(ctypes.Structure):
    _fields_ = [("inputDS", ctypes.c_char_p), ("wtsDS", ctypes.c_char_p), ("outputDS", ctypes.c_char_p),
                ("blocksize", ctypes.c_int), ("memsize", ctypes.c_int), ("band1", ctypes.c_int),
                ("band2", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int),
                ("computeCoherence", ctypes.c_int), ("verbose", ctypes.c_int)]


class CalampOptions(ctypes.Structure):
    _fields_ = [("inputDS", ctypes.c_char_p), ("maskDS", ctypes.c_char_p), ("outputDS", ctypes.c_char_p),
                ("defaultValue", ctypes.c_double), ("applySqrt", ctypes.c_int), ("blocksize", ctypes.c_int),
                ("memsize", ctypes.c_int), ("verbose", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    _fields_ = [("tile", LinkTile), ("d_slc", ctypes.c_void_p), ("d_wts", ctypes.c_void_p),
                ("d_phase", ctypes.c_void_p), ("d_tcorr", ctypes.c_void_p), ("d_comp", ctypes.c_void_p)]


FRINGE_SCHED_GRAM_VARIANT, FRINGE_SCHED_SOLVE, FRINGE_SCHED_GENERIC_A = 0, 1, 2
GRAM_VARIANTS = {"default": 0, "staged12": 1, "staged8": 2, "staged4": 3, "unstaged": 4}
# FRINGE_SCHED_GENERIC_A values: above 32 dates the solver's rows in
# registers (up to 128 dates), or one row per lane (up to 64 dates), or its
# work matrix in LDS or in global memory (above 64 dates)
GENERIC_A_PLACES = {"lanes": 0, "lds": 0, "global": 1, "registers": 2}
SCHEDULE_FILE = os.path.join(_HERE, "schedules.json")


def load_schedules(L, path=None):
    """Install the autotuned launch schedules (fringe_amd/autotune.py writes
    them) into the library's table.  FRINGE_SCHEDULES=<file> picks another
    table, FRINGE_SCHEDULES=none keeps the built-in defaults.  Returns the
    number of entries installed."""
    import json
    path = path or os.environ.get("FRINGE_SCHEDULES", SCHEDULE_FILE)
    if path == "none" or not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    k = 0
    for e in table.get("gram", []):
        if L.fringe_schedule_set(FRINGE_SCHED_GRAM_VARIANT, e["n"], e["Nx"], e["Ny"], GRAM_VARIANTS[e["variant"]]) == 0:
            k += 1
    for e in table.get("solve", []):
        v = 0 if e["launch"] == "grid" else (int(e["waves_per_cu"]) << 8) | int(e["grab"])
        if L.fringe_schedule_set(FRINGE_SCHED_SOLVE, e["n"], e["mode"], e["precision"], v) == 0:
            k += 1
    for e in table.get("generic_a", []):
        v = GENERIC_A_PLACES[e["work_matrix"]]
        if L.fringe_schedule_set(FRINGE_SCHED_GENERIC_A, e["n"], e["precision"], 0, v) == 0:
            k += 1
    return k


class ServerRequest(ctypes.Structure):
    _fields_ = [("tile", LinkTile), ("d_slc", ctypes.c_void_p), ("d_wts", ctypes.c_void_p),
                ("d_mask", ctypes.c_void_p), ("prob", ctypes.c_double)]


_LIB = None


def lib():
    """Load libfringe_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fringe_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.fringe_link_workspace_bytes.argtypes = [ctypes.POINTER(LinkTile)]
    L.fringe_link_workspace_bytes.restype = sz
    L.fringe_link_run.argtypes = [ctypes.POINTER(LinkTile), vp, vp, vp, vp, vp, vp, sz, vp]
    L.fringe_link_run_noalloc.argtypes = [ctypes.POINTER(LinkTile), vp, vp, vp, vp, vp, vp]
    L.fringe_link_resolved_pixels.argtypes = [ctypes.POINTER(LinkTile), vp, sz]
    L.fringe_link_resolved_pixels.restype = ctypes.c_longlong
    L.fringe_coherence_workspace_bytes.argtypes = [ctypes.POINTER(LinkTile)]
    L.fringe_coherence_workspace_bytes.restype = sz
    L.fringe_coherence_run.argtypes = [ctypes.POINTER(LinkTile), vp, vp, vp, vp, vp, sz, vp]
    L.fringe_nmap_workspace_bytes.argtypes = [ctypes.POINTER(NmapTile)]
    L.fringe_nmap_workspace_bytes.restype = sz
    L.fringe_nmap_run.argtypes = [ctypes.POINTER(NmapTile), vp, vp, vp, vp, vp, vp, sz, vp]
    L.fringe_ks_table.argtypes = [ctypes.c_int, ctypes.c_double, vp]
    L.fringe_ks_table.restype = ctypes.c_double
    L.fringe_ad2_threshold.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    L.fringe_ad2_threshold.restype = ctypes.c_double
    L.fringe_mask_words.argtypes = [ctypes.c_int, ctypes.c_int]
    L.fringe_version.restype = ctypes.c_char_p
    L.fringe_last_error.restype = ctypes.c_char_p
    L.fringe_timing_enable.argtypes = [ctypes.c_int]
    L.fringe_timing_read.argtypes = [vp, vp, ctypes.c_int]
    L.fringe_timing_stage_name.argtypes = [ctypes.c_int]
    L.fringe_timing_stage_name.restype = ctypes.c_char_p
    L.fringe_raster_info.argtypes = [ctypes.c_char_p, vp, vp, vp]
    L.fringe_raster_metadata.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                         ctypes.c_char_p, ctypes.c_int]
    L.fringe_raster_read.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 6 + [vp]
    L.fringe_adjust_wrapped.argtypes = [vp, vp, vp, ctypes.c_longlong, vp]
    L.fringe_adjust_unwrapped.argtypes = [vp, vp, vp, ctypes.c_longlong, vp]
    rq = ctypes.POINTER(ServerRequest)
    ll = ctypes.c_longlong
    L.fringe_server_request_bytes.argtypes = [rq]
    L.fringe_server_request_bytes.restype = sz
    L.fringe_server_create.argtypes = [ctypes.c_int, sz, ctypes.c_int, ctypes.POINTER(vp)]
    L.fringe_server_submit.argtypes = [vp, rq, ctypes.POINTER(ll)]
    L.fringe_server_poll.argtypes = [vp, ll, ctypes.POINTER(ctypes.c_int)]
    L.fringe_server_wait.argtypes = [vp, ll]
    L.fringe_server_result.argtypes = [vp, ll] + [ctypes.POINTER(vp)] * 5
    L.fringe_server_times.argtypes = [vp, ll, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.fringe_server_release.argtypes = [vp, ll]
    L.fringe_server_destroy.argtypes = [vp]
    for name, st in (("fringe_evd_options_init", EvdOptions), ("fringe_evd_process", EvdOptions),
                     ("fringe_phase_link_process", EvdOptions),
                     ("fringe_nmap_options_init", NmapOptions), ("fringe_nmap_process", NmapOptions),
                     ("fringe_despeck_options_init", DespeckOptions),
                     ("fringe_despeck_process", DespeckOptions),
                     ("fringe_calamp_options_init", CalampOptions), ("fringe_calamp_process", CalampOptions)):
        if hasattr(L, name):
            getattr(L, name).argtypes = [ctypes.POINTER(st)]
    bi = ctypes.POINTER(BatchItem)
    L.fringe_link_batch_workspace_bytes.argtypes = [bi, ctypes.c_int]
    L.fringe_link_batch_workspace_bytes.restype = sz
    L.fringe_link_batch_run.argtypes = [bi, ctypes.c_int, vp, sz, vp]
    L.fringe_schedule_set.argtypes = [ctypes.c_int] * 5
    L.fringe_schedule_get.argtypes = [ctypes.c_int] * 4
    load_schedules(L)
    _LIB = L
    return L


def last_error():
    return lib().fringe_last_error().decode()


def stage_times():
    """{stage: (ms, calls)} accumulated since the last reset (waits on events)."""
    import numpy as np
    L = lib()
    ms = np.zeros(8, np.float64)
    calls = np.zeros(8, np.int64)
    k = L.fringe_timing_read(ms.ctypes.data, calls.ctypes.data, 8)
    return {L.fringe_timing_stage_name(i).decode(): (float(ms[i]), int(calls[i])) for i in range(k)}
