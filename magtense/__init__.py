"""``magtense`` with the reference's module layout, running on the MI355X.

``magtense.micromag`` (``MicromagProblem(...).run_simulation(...)``) and
``magtense.magstatics`` (``Tiles``, ``grid_config``, ``get_demag_tensor``,
``get_H_field``) keep the interface of python/src/magtense/micromag.py and
magstatics.py, so callers such as python/examples/micromagnetism/
std_problem_4.py import them unchanged.  The computation runs in the
hand-written gfx950 HIP library of ``magtense_amd`` (``libmagtense_hip.so``,
C ABI in include/magtense_hip.h, loaded with ctypes); there is no CPU
fallback.  The reference's plotting helpers (``magtense.utils``), the
Halbach-cylinder builder and the iterative soft-magnet solver are not part of
this package.
"""

from . import micromag  # noqa: F401
from . import magstatics  # noqa: F401

__version__ = "0.1.0"
