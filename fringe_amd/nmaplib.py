"""nmaplib.Nmap: options and run() of the nmap module
(src/nmap/nmaplib.pyx:27-135), executed by fringe_nmap_process on the GPU."""
from ._options import _NmapOptions


class Nmap(_NmapOptions):
    pass
