"""HIP-backed counterparts of ``leopard_em.backend`` (``src/leopard_em/backend/__init__.py:3-12``)."""

from .core_match_template import MatchPlan, core_match_template, raw_statistics  # noqa: F401

__all__ = ["core_match_template", "MatchPlan", "raw_statistics"]
