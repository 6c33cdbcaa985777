"""MI355X-native (gfx950) probabilistic IVP solver path of probdiffeq.

``capi`` binds the C ABI of the HIP library; ``ivpsolve`` and ``probdiffeq``
mirror the reference's solver API on top of it.
"""

from probdiffeq_amd import capi  # noqa: F401

__all__ = ["capi"]
