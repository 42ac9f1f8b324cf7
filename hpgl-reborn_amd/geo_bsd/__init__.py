"""geo_bsd -- HPGL's kriging API (ordinary, simple and median indicator
kriging on SugarboxGrid properties) computed by hand-written HIP kernels on
an AMD Instinct MI355X (gfx950).

Import-compatible with the reference package's kriging entry points:
``from geo_bsd import *`` provides ordinary_kriging, simple_kriging,
lvm_kriging, median_ik, indicator_kriging, simple_kriging_weights,
sgs_simulation (sequential Gaussian simulation), CdfData, calc_cdf,
SugarboxGrid, CovarianceModel, ContProperty, IndProperty and covariance.  ``geo_bsd.device`` adds the
device-tensor entry points (variance output, mixed precision, multi-threshold
indicator CDFs).
"""
from .geo import *  # noqa: F401,F403
from . import validation  # noqa: F401
from .cdf import CdfData, calc_cdf  # noqa: F401
from .sgs import sgs_simulation  # noqa: F401
from . import cvariogram  # noqa: F401  (geo_bsd.cvariogram after "import geo_bsd", as the reference)
