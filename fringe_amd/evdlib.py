"""evdlib.Evd: options and run() of the evd module (src/evd/evdlib.cpp:14-41),
executed by fringe_evd_process on the GPU."""
from ._options import _LinkOptions


class Evd(_LinkOptions):
    _entry = "fringe_evd_process"
