"""despecklib.Despeck: options and run() of the despeck module
(src/despeck/despecklib.pyx:26-125), executed by fringe_despeck_process on the
GPU."""
from ._options import _DespeckOptions


class Despeck(_DespeckOptions):
    pass
