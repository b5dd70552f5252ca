"""phase_linklib.Phaselink: options and run() of the phase_link module
(src/phase_link/phase_linklib.pyx:33-157), executed by
fringe_phase_link_process on the GPU."""
from ._options import _LinkOptions


class Phaselink(_LinkOptions):
    _entry = "fringe_phase_link_process"
