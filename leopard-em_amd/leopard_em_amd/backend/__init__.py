"""HIP-backed counterparts of ``leopard_em.backend`` (``src/leopard_em/backend/__init__.py:3-12``)."""

from .core_match_template import MatchPlan, core_match_template, raw_statistics  # noqa: F401
from .core_refine_template import core_refine_template, refine_template_ragged  # noqa: F401

__all__ = ["core_match_template", "core_refine_template", "refine_template_ragged", "MatchPlan", "raw_statistics"]
