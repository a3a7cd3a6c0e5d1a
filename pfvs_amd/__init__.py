"""pfvs_amd: the coniZpM search of pfvs on MI355X (gfx950).

The public names mirror the reference package's coniZpM path
(pfvs/__init__.py): coniZpM, coniZp (every PFV of a direction, through its
dilation bound in pfvs_amd.dilation), CYData, IncompleteSearchError, the
gpu module and the p-vector enumeration.
"""

from . import gpu  # noqa: F401
from . import dilation  # noqa: F401
from .coni import coniZp, coniZpM  # noqa: F401
from .cydata import CYData  # noqa: F401
from .pvectors import box_enum, pvecs  # noqa: F401
from .util import IncompleteSearchError  # noqa: F401
