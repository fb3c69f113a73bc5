"""Peak extraction from search results: valid crop, z-score cutoff, greedy picking.

Counterparts of ``analysis/zscore_metric.py:14-166``, ``analysis/utils.py:6-48``,
``analysis/match_template_peaks.py`` and ``MatchTemplateResult.apply_valid_cropping``
(``pydantic_models/results/match_template_result.py:206-236``), same names and
argument meaning.  These run once per micrograph on the output maps (host side, as in
the reference).  Differences:

* the greedy distance suppression visits only the picked peaks in the neighbouring
  cells of a grid of cell size ceil(radius), instead of measuring every candidate
  against every picked peak; distances are compared in float32 like the
  reference's ``torch.norm``, so the picked set is the same;
* candidates with equal z-scores are taken in row-major pixel order (a stable
  sort; the reference's ``torch.sort`` leaves the order of ties unspecified).
"""

from __future__ import annotations

import math
import warnings
from typing import NamedTuple, Optional

import numpy as np
import pandas as pd
import torch
from scipy.special import erfcinv  # pylint: disable=no-name-in-module


class MatchTemplatePeaks(NamedTuple):
    """``analysis/match_template_peaks.py``: peak positions and their statistics."""

    pos_y: torch.Tensor
    pos_x: torch.Tensor
    mip: torch.Tensor
    scaled_mip: torch.Tensor
    psi: torch.Tensor
    theta: torch.Tensor
    phi: torch.Tensor
    relative_defocus: torch.Tensor
    correlation_mean: torch.Tensor
    correlation_variance: torch.Tensor
    total_correlations: int


def match_template_peaks_to_dict(peaks: MatchTemplatePeaks) -> dict:
    return peaks._asdict()


def match_template_peaks_to_dataframe(peaks: MatchTemplatePeaks) -> pd.DataFrame:
    return pd.DataFrame(peaks._asdict())


def valid_crop(maps: torch.Tensor, template_shape: tuple[int, int]) -> torch.Tensor:
    """[..., 0:H-h+1, 0:W-w+1]: positions where the template lies inside the image."""
    img_h, img_w = maps.shape[-2:]
    h, w = template_shape
    return maps[..., : img_h - h + 1, : img_w - w + 1]


def gaussian_noise_zscore_cutoff(num_ccg: int, false_positives: float = 1.0) -> float:
    """z such that a standard normal exceeds it ``false_positives`` times in ``num_ccg``
    draws: sqrt(2) * erfcinv(2 fp / num_ccg)  (``zscore_metric.py:14-39``)."""
    return float(erfcinv(2.0 * false_positives / num_ccg) * np.sqrt(2.0))


def filter_peaks_by_distance(
    peak_values: torch.Tensor,
    peak_locations: torch.Tensor,
    distance_threshold: float,
) -> torch.Tensor:
    """Greedy suppression in descending value order (``analysis/utils.py:6-48``).

    A candidate is dropped when an already picked peak lies closer than
    ``distance_threshold`` pixels.  Returns the picked (y, x) as int64 (n, 2).
    """
    vals = peak_values.detach().cpu().numpy()
    locs = peak_locations.detach().cpu().numpy().astype(np.int64)
    if locs.shape[0] == 0:
        return torch.zeros((0, 2), dtype=torch.long)
    order = np.argsort(-vals, kind="stable")
    locs = locs[order]
    thr = np.float32(distance_threshold)
    cell = max(1, int(math.ceil(float(distance_threshold))))
    grid: dict[tuple[int, int], list[int]] = {}
    picked: list[int] = []
    for i in range(locs.shape[0]):
        y, x = int(locs[i, 0]), int(locs[i, 1])
        cy, cx = y // cell, x // cell
        suppressed = False
        for gy in (cy - 1, cy, cy + 1):
            for gx in (cx - 1, cx, cx + 1):
                for j in grid.get((gy, gx), ()):
                    dy = np.float32(y - locs[j, 0])
                    dx = np.float32(x - locs[j, 1])
                    if np.sqrt(dy * dy + dx * dx, dtype=np.float32) < thr:
                        suppressed = True
                        break
                if suppressed:
                    break
            if suppressed:
                break
        if suppressed:
            continue
        picked.append(i)
        grid.setdefault((cy, cx), []).append(i)
    return torch.from_numpy(locs[picked].copy())


def find_peaks_from_zscore(
    zscore_map: torch.Tensor,
    zscore_cutoff: float,
    mask_radius: float = 5.0,
) -> tuple[torch.Tensor, torch.Tensor]:
    """(pos_y, pos_x) of the picked pixels above the cutoff (``zscore_metric.py:42-83``)."""
    z = zscore_map.detach()
    above = torch.nonzero(z > zscore_cutoff, as_tuple=False)
    if above.shape[0] == 0:
        return torch.tensor([], dtype=torch.long), torch.tensor([], dtype=torch.long)
    vals = z[above[:, 0], above[:, 1]]
    picked = filter_peaks_by_distance(vals, above, mask_radius)
    return picked[:, 0], picked[:, 1]


# pylint: disable=too-many-arguments,too-many-positional-arguments
def extract_peaks_and_statistics_zscore(
    mip: torch.Tensor,
    scaled_mip: torch.Tensor,
    best_psi: torch.Tensor,
    best_theta: torch.Tensor,
    best_phi: torch.Tensor,
    best_defocus: torch.Tensor,
    correlation_average: torch.Tensor,
    correlation_variance: torch.Tensor,
    total_correlation_positions: int,
    false_positives: float = 1.0,
    z_score_cutoff: Optional[float] = None,
    mask_radius: float = 5.0,
) -> MatchTemplatePeaks:
    """Peaks of ``scaled_mip`` and the maps' values there (``zscore_metric.py:89-166``).

    The default cutoff allows ``false_positives`` false peaks among
    mip.numel() * total_correlation_positions Gaussian-noise samples.
    """
    if z_score_cutoff is None:
        z_score_cutoff = gaussian_noise_zscore_cutoff(
            num_ccg=mip.numel() * total_correlation_positions, false_positives=false_positives)
    pos_y, pos_x = find_peaks_from_zscore(scaled_mip, z_score_cutoff, mask_radius)
    if len(pos_y) == 0:
        warnings.warn("No peaks found using z-score metric.", stacklevel=2)
    return MatchTemplatePeaks(
        pos_y=pos_y,
        pos_x=pos_x,
        mip=mip[pos_y, pos_x],
        scaled_mip=scaled_mip[pos_y, pos_x],
        psi=best_psi[pos_y, pos_x],
        theta=best_theta[pos_y, pos_x],
        phi=best_phi[pos_y, pos_x],
        relative_defocus=best_defocus[pos_y, pos_x],
        correlation_mean=correlation_average[pos_y, pos_x],
        correlation_variance=correlation_variance[pos_y, pos_x],
        total_correlations=total_correlation_positions,
    )
