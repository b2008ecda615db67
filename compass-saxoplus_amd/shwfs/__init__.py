"""Host side of the MI355X Shack-Hartmann WFS path.

The native modules ``carma`` and ``sutra`` (built in-tree next to this
package by ``build.py``) carry the reference's class API; this package holds
what the reference computes in Python before it calls them:

* ``params``   - the parameter fields of shesha's ParamTel/ParamGeom/ParamWfs/
                 ParamCentroider that the SH path reads
* ``geom``     - pupil and SH geometry tables (shesha/shesha/init/geom_init.py)
* ``wfs_init`` - Telescope / Sensors construction and table upload
                 (shesha/shesha/init/wfs_init.py)
* ``rtc_init`` - centroider construction and WCoG weights
                 (shesha/shesha/init/rtc_init.py:272-433)
* ``screens``  - seeded von Karman phase screens (synthetic inputs)
* ``configs``  - the five named workloads of this path
* ``system``   - a one-WFS convenience wrapper used by tests and the benchmark
"""
import os
import sys

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG_ROOT not in sys.path:
    sys.path.insert(0, _PKG_ROOT)


def native():
    """Imports the native modules; raises if the HIP build is missing."""
    try:
        import carma  # noqa: F401
        import sutra  # noqa: F401
    except ImportError as e:  # no CPU fallback exists: the path is HIP only
        raise ImportError(
            "the native sutra/carma modules are not built; run "
            "`python compass-saxoplus_amd/build.py` (" + str(e) + ")") from e
    return carma, sutra
