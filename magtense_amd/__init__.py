"""Native layer of the MI355X MagTense micromagnetic path.

``magtense_amd._lib`` loads ``libmagtense_hip.so`` (hand-written gfx950 HIP
kernels behind the C ABI declared in include/magtense_hip.h, built in-tree
from ``magtense_amd/csrc``); ``magtense_amd.adjoint`` is the differentiable
RK4 integration (torch.autograd).  The reference's Python API lives in the
``magtense`` package (``magtense.micromag``, ``magtense.magstatics``).
"""

from . import _lib  # noqa: F401

__version__ = "0.1.0"
