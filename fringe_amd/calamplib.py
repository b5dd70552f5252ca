"""calamplib.Calamp: options and run() of the calamp module
(src/calamp/calamplib.pyx:21-100), executed by fringe_calamp_process on the
GPU."""
from ._options import _CalampOptions


class Calamp(_CalampOptions):
    pass
