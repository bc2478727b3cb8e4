"""MI355X-native q-LSKUM pseudo-time iteration (3-D meshfree kinetic Euler).

Drop-in for the fpi_solver hot path of Pathlessbark8/3d-mfcfd: the same
entry points over the same point and connectivity arrays, computed by
hand-written CDNA4 HIP kernels in ``libmfcfd_hip.so`` behind a C ABI
(``include/mfcfd_hip.h``).

The directory name is not a Python identifier; import it with
``importlib.import_module("3d-mfcfd_amd")``.  The alias ``mfcfd_amd`` is
registered on import.
"""

import sys as _sys

from . import pointcloud  # noqa: F401
from . import solver  # noqa: F401
from . import postprocess  # noqa: F401

_sys.modules.setdefault("mfcfd_amd", _sys.modules[__name__])
