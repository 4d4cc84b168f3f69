"""MI355X-native rebuild of MagTense's micromagnetic hot path.

Drop-in for ``magtense.micromag`` (``MicromagProblem(...).run_simulation(...)``)
backed by hand-written gfx950 HIP kernels behind a C ABI
(``include/magtense_hip.h``, loaded with ctypes).  ``import magtense_amd as
magtense`` gives ``magtense.micromag.MicromagProblem``.
"""

from . import micromag  # noqa: F401
from . import _lib  # noqa: F401

__version__ = "0.1.0"
