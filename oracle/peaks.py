"""z-score cutoff and greedy peak picking (CPU oracle, test-only).

Restates ``analysis/zscore_metric.py:14-83`` and ``analysis/utils.py:6-48``.
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.special import erfcinv


def gaussian_noise_zscore_cutoff(num_ccg: int, false_positives: float = 1.0) -> float:
    """``analysis/zscore_metric.py:14-39``: sqrt(2) * erfcinv(2 fp / num_ccg)."""
    return float(erfcinv(2.0 * false_positives / num_ccg) * np.sqrt(2.0))


def filter_peaks_by_distance(peak_values: torch.Tensor, peak_locations: torch.Tensor,
                             distance_threshold: float) -> torch.Tensor:
    """``analysis/utils.py:6-48``: descending sort, suppress peaks closer than the radius."""
    peak_values, order = torch.sort(peak_values, descending=True, stable=True)
    loc = peak_locations[order]
    taken = torch.zeros(loc.shape[0], dtype=torch.bool)
    picked = []
    for i in range(loc.shape[0]):
        if taken[i]:
            continue
        picked.append(loc[i])
        d = torch.norm(loc.to(torch.float32) - loc[i].to(torch.float32), dim=1)
        taken |= d < distance_threshold
    if not picked:
        return torch.zeros((0, 2), dtype=torch.long)
    return torch.stack(picked, 0)


def find_peaks_from_zscore(zscore_map: torch.Tensor, cutoff: float, mask_radius: float = 5.0):
    """``analysis/zscore_metric.py:42-83`` -> (pos_y, pos_x)."""
    peaks = torch.nonzero(zscore_map > cutoff, as_tuple=False)
    if peaks.shape[0] == 0:
        return torch.tensor([], dtype=torch.long), torch.tensor([], dtype=torch.long)
    vals = zscore_map[tuple(peaks.t())]
    picked = filter_peaks_by_distance(vals, peaks, mask_radius)
    return picked[:, 0], picked[:, 1]
