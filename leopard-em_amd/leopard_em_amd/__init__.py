"""leopard_em_amd -- MI355X-native backend for Leopard-EM's 2DTM search path.

Drop-in counterparts of ``leopard_em.backend.core_match_template`` and
``core_refine_template`` running on hand-written gfx950 HIP kernels through the
C ABI in ``include/tdtm.h``.
"""

from .backend import core_match_template, core_refine_template  # noqa: F401

__all__ = ["core_match_template", "core_refine_template"]
