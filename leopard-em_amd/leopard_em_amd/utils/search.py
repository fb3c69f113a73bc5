"""Search grids: orientations, relative defocus and pixel-size values.

* ``get_uniform_euler_angles``: the 'uniform' base grid used by
  ``OrientationSearchConfig.euler_angles`` (``config/orientation_search.py:127-170``,
  torch_so3's ``get_uniform_euler_angles``): theta rings every ``theta_step`` degrees
  from theta_min to theta_max inclusive, max(1, round(|sin theta| * phi_range /
  theta_step)) evenly spaced phi values per ring, psi every ``psi_step`` in
  [psi_min, psi_max).  Rows are (phi, theta, psi) in degrees, ring-major then phi then
  psi.
* ``get_search_tensors``: ``utils/search_utils.py:57-93``.
* ``get_cs_range``: ``utils/search_utils.py:6-29``.
"""

from __future__ import annotations

import numpy as np
import torch


def get_uniform_euler_angles(
    psi_step: float,
    theta_step: float,
    phi_min: float = 0.0,
    phi_max: float = 360.0,
    theta_min: float = 0.0,
    theta_max: float = 180.0,
    psi_min: float = 0.0,
    psi_max: float = 360.0,
) -> torch.Tensor:
    """(N, 3) float32 (phi, theta, psi) degrees."""
    if psi_step <= 0 or theta_step <= 0:
        raise ValueError("psi_step and theta_step must be positive")
    n_theta = int(np.floor((theta_max - theta_min) / theta_step + 0.5 + 1e-9)) + 1
    thetas = theta_min + theta_step * np.arange(n_theta, dtype=np.float64)
    thetas = thetas[thetas <= theta_max + 1e-9]
    n_psi = max(1, int(round((psi_max - psi_min) / psi_step)))
    psis = psi_min + psi_step * np.arange(n_psi, dtype=np.float64)
    phi_range = phi_max - phi_min
    blocks = []
    for th in thetas:
        n_phi = max(1, int(round(abs(np.sin(np.deg2rad(th))) * phi_range / theta_step)))
        phis = phi_min + (phi_range / n_phi) * np.arange(n_phi, dtype=np.float64)
        blk = np.empty((n_phi, n_psi, 3), dtype=np.float64)
        blk[:, :, 0] = phis[:, None]
        blk[:, :, 1] = th
        blk[:, :, 2] = psis[None, :]
        blocks.append(blk.reshape(-1, 3))
    return torch.from_numpy(np.concatenate(blocks, 0).astype(np.float32))


def get_search_tensors(min_val: float, max_val: float, step_size: float,
                       skip_enforce_zero: bool = False) -> torch.Tensor:
    """arange(min, max + step, step) (float32), with 0 inserted in sorted position if
    no value is within 1e-6 of it."""
    vals = torch.arange(min_val, max_val + step_size, step_size, dtype=torch.float32)
    if not skip_enforce_zero and float(torch.min(torch.abs(vals))) > 1e-6:
        vals = torch.sort(torch.cat([vals, torch.zeros(1, dtype=torch.float32)])).values
    return vals


def get_cs_range(pixel_size: float, pixel_size_offsets: torch.Tensor, cs: float = 2.7) -> torch.Tensor:
    """Spherical aberration that keeps Cs / pixel^4 fixed: cs / ((p + dp) / p)^4."""
    return cs / torch.pow((pixel_size + pixel_size_offsets) / pixel_size, 4)
