"""brian_amd: MI355X (gfx950) backend for Brian 1.4.4's clock-driven network step.

``from brian_amd import *`` gives the Brian 1 names a network script uses
(NeuronGroup, Connection, DelayConnection, ExponentialSTDP, SpikeMonitor,
PopulationRateMonitor, Network, run, units, ...).  The step itself runs in
libbrianmi.so on the GPU; there is no CPU simulation path.
"""
__version__ = "0.1.0"

import numpy as _np
from numpy import *  # noqa: F401,F403  (Brian scripts rely on `from brian import *` giving numpy)

from .units import *  # noqa: F401,F403
from .units import unit_namespace  # noqa: F401
from .equations import Equations  # noqa: F401
from .core import *  # noqa: F401,F403
from .statistics import CV, firing_rate, gamma_factor, get_gamma_factor, NaN  # noqa: F401
from .spikeio import load_aer, load_raster, save_raster  # noqa: F401
from numpy.random import rand, randn, seed  # noqa: F401
