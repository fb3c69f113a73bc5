"""Rotations, Fourier central slices and search grids (CPU oracle, test-only).

Restates the third-party geometry the reference calls:
* ``roma.euler_to_rotmat("ZYZ", deg)`` -- ``backend/core_match_template.py:506-508``
* ``roma.rotmat_to_euler`` / composition -- ``backend/utils.py:25-54``
* ``torch_fourier_slice.extract_central_slices_rfft_3d`` --
  ``backend/cross_correlation.py:76-79, 225-228, 330-333``
* ``torch_fourier_slice.transform_slice_2d`` (anisotropic magnification) --
  ``backend/cross_correlation.py:81-88, 231-238, 336-343``
* ``torch_so3.get_uniform_euler_angles`` -- ``config/orientation_search.py:160-170``
* ``get_search_tensors`` / ``get_cs_range`` -- ``utils/search_utils.py:6-93``
"""

from __future__ import annotations

import math

import torch


# --------------------------------------------------------------------------- rotations
def _rz(a: torch.Tensor) -> torch.Tensor:
    c, s = torch.cos(a), torch.sin(a)
    z, o = torch.zeros_like(a), torch.ones_like(a)
    return torch.stack(
        [torch.stack([c, -s, z], -1), torch.stack([s, c, z], -1), torch.stack([z, z, o], -1)],
        -2,
    )


def _ry(a: torch.Tensor) -> torch.Tensor:
    c, s = torch.cos(a), torch.sin(a)
    z, o = torch.zeros_like(a), torch.ones_like(a)
    return torch.stack(
        [torch.stack([c, z, s], -1), torch.stack([z, o, z], -1), torch.stack([-s, z, c], -1)],
        -2,
    )


def euler_to_rotmat_zyz(angles_deg: torch.Tensor, dtype=torch.float64) -> torch.Tensor:
    """Intrinsic ZYZ Euler angles (phi, theta, psi) in degrees -> rotation matrices.

    R = Rz(phi) @ Ry(theta) @ Rz(psi): roma's upper-case (intrinsic) convention as used
    at ``backend/core_match_template.py:506-508`` and
    ``backend/core_refine_template.py:644-648``.
    """
    a = torch.deg2rad(angles_deg.to(dtype))
    return _rz(a[..., 0]) @ _ry(a[..., 1]) @ _rz(a[..., 2])


def rotmat_to_euler_zyz(rot: torch.Tensor) -> torch.Tensor:
    """Inverse of :func:`euler_to_rotmat_zyz`; degrees, theta in [0, 180].

    Closed form for R = Rz(phi) Ry(theta) Rz(psi):
    R22 = cos(theta), (R02, R12) = sin(theta) (cos phi, sin phi),
    (R20, R21) = sin(theta) (-cos psi, sin psi).  Follows the use of
    ``roma.rotmat_to_euler`` in ``backend/utils.py:52``.  At theta = 0 or 180 only
    phi +/- psi is defined; this restatement then puts the whole angle into phi.
    """
    r22 = rot[..., 2, 2].clamp(-1.0, 1.0)
    theta = torch.arccos(r22)
    st = torch.sin(theta)
    gimbal = st.abs() < 1e-7
    phi = torch.atan2(rot[..., 1, 2], rot[..., 0, 2])
    psi = torch.atan2(rot[..., 2, 1], -rot[..., 2, 0])
    # gimbal lock: R = Rz(phi +/- psi) (theta = 0) -> psi = 0
    phi_g = torch.atan2(rot[..., 1, 0], rot[..., 0, 0])
    phi_g = torch.where(r22 < 0, torch.atan2(-rot[..., 1, 0], -rot[..., 0, 0]), phi_g)
    phi = torch.where(gimbal, phi_g, phi)
    psi = torch.where(gimbal, torch.zeros_like(psi), psi)
    return torch.rad2deg(torch.stack([phi, theta, psi], -1))


def combine_euler_angles(angle_a: torch.Tensor, angle_b: torch.Tensor) -> torch.Tensor:
    """Compose two ZYZ rotations, R = R(a) @ R(b) -> Euler; ``backend/utils.py:25-54``."""
    ra = euler_to_rotmat_zyz(angle_a)
    rb = euler_to_rotmat_zyz(angle_b)
    return rotmat_to_euler_zyz(ra @ rb).to(angle_a.dtype)


# ---------------------------------------------------------------- central slice (K1/K2)
def central_slice_fftfreq_grid(n: int, dtype=torch.float64) -> torch.Tensor:
    """(n, n//2+1, 3) grid of (x, y, z=0) fftfreqs for an fftshifted-in-y rfft slice.

    Row index iy <-> fy = (iy - n//2)/n, column ix <-> fx = ix/n
    (torch_fourier_slice's ``_central_slice_fftfreq_grid(rfft=True, fftshift=True)``).
    """
    fy = (torch.arange(n, dtype=dtype) - n // 2) / n
    fx = torch.arange(n // 2 + 1, dtype=dtype) / n
    gy, gx = torch.meshgrid(fy, fx, indexing="ij")
    return torch.stack([gx, gy, torch.zeros_like(gx)], -1)


def trilinear_sample_zero_outside(volume: torch.Tensor, coords_zyx: torch.Tensor) -> torch.Tensor:
    """Trilinear samples of a (d, h, w) array at array coordinates (..., 3) in zyx order.

    Samples with any coordinate outside [0, n-1] are zero; neighbours beyond the last
    index are clamped (``grid_sample(align_corners=True, padding_mode='border')``
    followed by zeroing of outside points, as torch_image_interpolation does).
    """
    shape = torch.tensor(volume.shape, dtype=coords_zyx.dtype)
    inside = ((coords_zyx >= 0) & (coords_zyx <= shape - 1)).all(-1)
    c = torch.minimum(torch.maximum(coords_zyx, torch.zeros_like(coords_zyx)), shape - 1)
    i0 = torch.floor(c).long()
    frac = c - i0.to(c.dtype)
    i1 = torch.minimum(i0 + 1, (shape - 1).long())
    out = torch.zeros(coords_zyx.shape[:-1], dtype=volume.dtype)
    for dz in (0, 1):
        iz = i1[..., 0] if dz else i0[..., 0]
        wz = frac[..., 0] if dz else 1 - frac[..., 0]
        for dy in (0, 1):
            iy = i1[..., 1] if dy else i0[..., 1]
            wy = frac[..., 1] if dy else 1 - frac[..., 1]
            for dx in (0, 1):
                ix = i1[..., 2] if dx else i0[..., 2]
                wx = frac[..., 2] if dx else 1 - frac[..., 2]
                out = out + (wz * wy * wx).to(volume.dtype) * volume[iz, iy, ix]
    return torch.where(inside, out, torch.zeros_like(out))


def extract_central_slices_rfft_3d(volume_rfft: torch.Tensor, rotation_matrices: torch.Tensor) -> torch.Tensor:
    """Central slices of an fftshifted (dims 0,1) 3D rfft, one per rotation matrix.

    Follows torch_fourier_slice ``extract_central_slices_rfft_3d`` as called at
    ``backend/cross_correlation.py:76-79``: the slice-plane frequencies (x, y, 0) are
    rotated by R (R acts on xyz column vectors), points that land at x < 0 are
    negated and their samples conjugated (Hermitian fold), frequencies map to array
    coordinates ``n//2 + f*n`` (dims 0, 1) and ``f*n`` (rfft dim), and the volume is
    sampled trilinearly with zero outside.  Output (..., n, n//2+1), fftshifted in y.
    """
    n = volume_rfft.shape[0]
    real_dtype = torch.float64 if volume_rfft.dtype == torch.complex128 else torch.float32
    grid = central_slice_fftfreq_grid(n, real_dtype)  # (n, n/2+1, 3) xyz
    rot = rotation_matrices.to(real_dtype)
    rotated = torch.einsum("...ij,hwj->...hwi", rot, grid)  # xyz
    flip = rotated[..., 0] < 0
    rotated = torch.where(flip[..., None], -rotated, rotated)
    zyx = torch.stack(
        [rotated[..., 2] * n + n // 2, rotated[..., 1] * n + n // 2, rotated[..., 0] * n], -1
    )
    samples = trilinear_sample_zero_outside(volume_rfft, zyx)
    return torch.where(flip, samples.conj(), samples)



def bilinear_sample_zero_outside(image: torch.Tensor, coords_yx: torch.Tensor) -> torch.Tensor:
    """Bilinear samples of an (h, w) array at array coordinates (..., 2) in yx order.

    Same edge rule as :func:`trilinear_sample_zero_outside`: samples with a coordinate
    outside [0, n-1] are zero, neighbours beyond the last index are clamped.
    """
    shape = torch.tensor(image.shape, dtype=coords_yx.dtype)
    inside = ((coords_yx >= 0) & (coords_yx <= shape - 1)).all(-1)
    c = torch.minimum(torch.maximum(coords_yx, torch.zeros_like(coords_yx)), shape - 1)
    i0 = torch.floor(c).long()
    frac = c - i0.to(c.dtype)
    i1 = torch.minimum(i0 + 1, (shape - 1).long())
    out = torch.zeros(coords_yx.shape[:-1], dtype=image.dtype)
    for dy in (0, 1):
        iy = i1[..., 0] if dy else i0[..., 0]
        wy = frac[..., 0] if dy else 1 - frac[..., 0]
        for dx in (0, 1):
            ix = i1[..., 1] if dx else i0[..., 1]
            wx = frac[..., 1] if dx else 1 - frac[..., 1]
            out = out + (wy * wx).to(image.dtype) * image[iy, ix]
    return torch.where(inside, out, torch.zeros_like(out))


def transform_slice_2d(slices: torch.Tensor, transform_matrix: torch.Tensor) -> torch.Tensor:
    """Anisotropic magnification of central slices (fftshifted-in-y rfft layout).

    Follows torch_fourier_slice ``transform_slice_2d`` as called at
    ``backend/cross_correlation.py:81-88`` (before the ifftshift / DC / contrast flip of
    ``:90-92``): every output frequency k = (x, y) of the (n, n//2+1) slice grid takes
    the input slice's value at k' = A k (A = ``transform_matrix``, 2 x 2, acting on
    (x, y) column vectors, the row / column order of ``OpticsGroup.mag_matrix_tensor``);
    points with x' < 0 are negated and their samples conjugated (Hermitian fold), and
    the slice is sampled bilinearly at array coordinates (n//2 + y' n, x' n) with zero
    outside.  Which of A or its inverse the upstream package applies, and its order of
    (x, y), are not pinned by any file of the reference (the package is absent): this
    restatement fixes k' = A k.
    """
    n = slices.shape[-2]
    real_dtype = torch.float64 if slices.dtype == torch.complex128 else torch.float32
    a = torch.as_tensor(transform_matrix, dtype=real_dtype).reshape(2, 2)
    g = central_slice_fftfreq_grid(n, real_dtype)[..., :2]  # (n, n/2+1, 2) xy
    t = torch.einsum("ij,hwj->hwi", a, g)
    flip = t[..., 0] < 0
    t = torch.where(flip[..., None], -t, t)
    yx = torch.stack([t[..., 1] * n + n // 2, t[..., 0] * n], -1)
    flat = slices.reshape(-1, n, n // 2 + 1)
    out = torch.stack([bilinear_sample_zero_outside(flat[i], yx) for i in range(flat.shape[0])], 0)
    out = out.reshape(slices.shape)
    return torch.where(flip, out.conj(), out)


# --------------------------------------------------------------------------- grids
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
    """Uniform (phi, theta, psi) grid, float32 degrees, shape (N, 3).

    Restates the 'uniform' base grid of torch_so3 ``get_uniform_euler_angles``
    (``config/orientation_search.py:160-170``): theta rings every ``theta_step`` from
    theta_min to theta_max inclusive, ``max(1, round(|sin theta| * phi_range /
    theta_step))`` phi values per ring, and psi every ``psi_step`` in [psi_min, psi_max).
    Ordering (phi-ring major, then psi) and the rounding rule are this restatement's
    choice: the orientation *count* is pinned by the reference test
    (``tests/pydantic_models/test_orientation_search.py:106-110``: 24 at 90/90 deg).
    """
    thetas = torch.arange(theta_min, theta_max + 0.5 * theta_step, theta_step, dtype=torch.float64)
    thetas = thetas[thetas <= theta_max + 1e-9]
    n_psi = max(1, int(round((psi_max - psi_min) / psi_step)))
    psis = psi_min + torch.arange(n_psi, dtype=torch.float64) * psi_step
    rows = []
    phi_range = phi_max - phi_min
    for th in thetas.tolist():
        n_phi = max(1, int(round(abs(math.sin(math.radians(th))) * phi_range / theta_step)))
        phis = phi_min + torch.arange(n_phi, dtype=torch.float64) * (phi_range / n_phi)
        for ph in phis.tolist():
            block = torch.empty((n_psi, 3), dtype=torch.float64)
            block[:, 0] = ph
            block[:, 1] = th
            block[:, 2] = psis
            rows.append(block)
    return torch.cat(rows, 0).to(torch.float32)


def get_search_tensors(min_val: float, max_val: float, step_size: float, skip_enforce_zero: bool = False) -> torch.Tensor:
    """``utils/search_utils.py:57-93``: arange(min, max+step, step) with 0 inserted."""
    vals = torch.arange(min_val, max_val + step_size, step_size, dtype=torch.float32)
    if abs(torch.min(torch.abs(vals))) > 1e-6 and not skip_enforce_zero:
        vals = torch.sort(torch.cat([vals, torch.tensor([0.0])]))[0]
    return vals


def get_cs_range(pixel_size: float, pixel_size_offsets: torch.Tensor, cs: float = 2.7) -> torch.Tensor:
    """``utils/search_utils.py:6-29``: Cs(p) = Cs0 / (p/p0)^4."""
    pixel_sizes = pixel_size + pixel_size_offsets
    return cs / torch.pow(pixel_sizes / pixel_size, 4)
