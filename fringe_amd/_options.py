"""Option objects mirroring the reference's Python bindings.

evdlib.Evd (src/evd/evdlib.cpp:14-41), phase_linklib.Phaselink
(src/phase_link/phase_linklib.pyx:33-157) and nmaplib.Nmap
(src/nmap/nmaplib.pyx:27-135) expose the module options as attributes and a
run() that calls the module processor.  Here run() calls the C ABI of
libfringe_amd.so; like the reference bindings it does not raise on a
module error, but it returns the module's return code (and keeps it in
``status``) instead of discarding it.
"""
import ctypes

from . import _lib

_PREC = {"mixed": _lib.FRINGE_PREC_MIXED, "fp64": _lib.FRINGE_PREC_FP64,
         "mixed_bf16": _lib.FRINGE_PREC_MIXED_BF16}


class _LinkOptions:
    # attribute name -> field of fringe_evd_options (evdlib.cpp:20-36)
    _map = {"inputDS": "inputDS", "weightsDS": "wtsDS", "outputFolder": "outputFolder",
            "outputCompressedSlcFolder": "outputCompressedSlcFolder", "compSlc": "compSlc",
            "minimumNeighbors": "minNeighbors", "miniStackCount": "miniStackCount",
            "blocksize": "blocksize", "memsize": "memsize", "halfWindowX": "Nx",
            "halfWindowY": "Ny", "method": "method", "bandWidth": "bandWidth"}
    _entry = None

    def __init__(self):
        o = _lib.EvdOptions()
        _lib.lib().fringe_evd_options_init(ctypes.byref(o))
        vals = {}
        for attr, field in self._map.items():
            v = getattr(o, field)
            vals[attr] = v.decode() if isinstance(v, bytes) else v
        vals["outputCompressedSlcFolder"] = ""
        object.__setattr__(self, "_vals", vals)
        object.__setattr__(self, "precision", "mixed")
        object.__setattr__(self, "verbose", True)
        object.__setattr__(self, "status", None)

    def __getattr__(self, name):
        vals = object.__getattribute__(self, "_vals")
        if name in vals:
            return vals[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self._map:
            self._vals[name] = value
        else:
            object.__setattr__(self, name, value)

    def _struct(self):
        o = _lib.EvdOptions()
        _lib.lib().fringe_evd_options_init(ctypes.byref(o))
        keep = []
        for attr, field in self._map.items():
            v = self._vals[attr]
            if isinstance(v, str):
                b = v.encode()
                keep.append(b)
                setattr(o, field, b)
            else:
                setattr(o, field, int(v))
        o.precision = _PREC[self.precision] if isinstance(self.precision, str) else int(self.precision)
        o.verbose = 1 if self.verbose else 0
        return o, keep

    def print(self):
        for attr in self._map:
            print(f"{attr} = {self._vals[attr]}")
        print(f"precision = {self.precision}")

    def run(self):
        o, _keep = self._struct()
        fn = getattr(_lib.lib(), self._entry)
        self.status = fn(ctypes.byref(o))
        return self.status


class _PlainOptions:
    """Options of a module whose C struct holds strings, ints and doubles;
    subclasses name the struct, its init and process entry points and the
    attribute -> field map of the reference binding."""
    _struct_cls = None
    _init = None
    _entry = None
    _map = {}
    _bool_attrs = ()

    def __init__(self):
        o = self._struct_cls()
        getattr(_lib.lib(), self._init)(ctypes.byref(o))
        vals = {}
        for attr, field in self._map.items():
            v = getattr(o, field)
            vals[attr] = v.decode() if isinstance(v, bytes) else v
        for attr in self._bool_attrs:
            vals[attr] = bool(vals[attr])
        object.__setattr__(self, "_vals", vals)
        object.__setattr__(self, "verbose", True)
        object.__setattr__(self, "status", None)

    def __getattr__(self, name):
        vals = object.__getattribute__(self, "_vals")
        if name in vals:
            return vals[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self._map:
            self._vals[name] = value
        else:
            object.__setattr__(self, name, value)

    def print(self):
        for attr in self._map:
            print(f"{attr} = {self._vals[attr]}")

    def run(self):
        o = self._struct_cls()
        getattr(_lib.lib(), self._init)(ctypes.byref(o))
        keep = []
        for attr, field in self._map.items():
            v = self._vals[attr]
            if isinstance(v, str):
                b = v.encode()
                keep.append(b)
                setattr(o, field, b)
            elif isinstance(v, float):
                setattr(o, field, v)
            else:
                setattr(o, field, int(v))
        o.verbose = 1 if self.verbose else 0
        self.status = getattr(_lib.lib(), self._entry)(ctypes.byref(o))
        return self.status


class _NmapOptions(_PlainOptions):
    # attribute name -> field of fringe_nmap_options (nmaplib.pyx:27-135)
    _struct_cls = _lib.NmapOptions
    _init = "fringe_nmap_options_init"
    _entry = "fringe_nmap_process"
    _map = {"inputDS": "inputDS", "countDS": "ncountDS", "weightsDS": "wtsDS", "maskDS": "maskDS",
            "method": "method", "minimumProbability": "prob", "blocksize": "blocksize",
            "memsize": "memsize", "halfWindowX": "Nx", "halfWindowY": "Ny", "noGPU": "noGPU"}
    _bool_attrs = ("noGPU",)


class _DespeckOptions(_PlainOptions):
    # attribute name -> field of fringe_despeck_options (despecklib.pyx:26-125)
    _struct_cls = _lib.DespeckOptions
    _init = "fringe_despeck_options_init"
    _entry = "fringe_despeck_process"
    _map = {"inputDS": "inputDS", "weightsDS": "wtsDS", "outputDS": "outputDS", "blocksize": "blocksize",
            "memsize": "memsize", "halfWindowX": "Nx", "halfWindowY": "Ny", "band1": "band1",
            "band2": "band2", "coherenceFlag": "computeCoherence"}
    _bool_attrs = ("coherenceFlag",)


class _CalampOptions(_PlainOptions):
    # attribute name -> field of fringe_calamp_options (calamplib.pyx:21-100)
    _struct_cls = _lib.CalampOptions
    _init = "fringe_calamp_options_init"
    _entry = "fringe_calamp_process"
    _map = {"inputDS": "inputDS", "outputDS": "outputDS", "maskDS": "maskDS", "defaultValue": "defaultValue",
            "applySqrt": "applySqrt", "blocksize": "blocksize", "memsize": "memsize"}
    _bool_attrs = ("applySqrt",)
