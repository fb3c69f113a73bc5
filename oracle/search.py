"""Projection normalisation, cross-correlation and search loops (CPU oracle, test-only).

Restates the reference backend in plain torch on the CPU:
* ``backend/utils.py:106-188``   normalize_template_projection (incl. the double
  relative-size scaling of the mean, ``:177-178``)
* ``backend/cross_correlation.py:287-369``  do_batched_orientation_cross_correlate_cpu
* ``backend/utils.py:191-256``   do_iteration_statistics_updates
* ``backend/process_results.py:7-183``  aggregate / decode / mean+"variance" / scale_mip
* ``backend/core_match_template.py:369-562`` single-device orientation loop
* ``backend/core_refine_template.py:555-767`` per-particle refinement loop,
  ``:771-895`` cross_correlate_particle_stack, ``:149-191`` / ``:504-536`` aggregation
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .filters import calculate_ctf_filter_stack_full_args
from .geometry import combine_euler_angles, euler_to_rotmat_zyz, extract_central_slices_rfft_3d, transform_slice_2d


def normalize_template_projection(projections: torch.Tensor, small_shape, large_shape) -> torch.Tensor:
    """``backend/utils.py:106-188``."""
    top = projections[..., 0, :]
    bottom = projections[..., -1, :]
    left = projections[..., 1:-1, 0]
    right = projections[..., 1:-1, -1]
    edge = torch.cat([top, bottom, left, right], dim=-1)
    projections = projections - edge.mean(dim=-1)[..., None, None]
    large_size_sqrt = (large_shape[0] * large_shape[1]) ** 0.5
    rel = (small_shape[0] * small_shape[1]) / (large_shape[0] * large_shape[1])
    mean = torch.mean(projections, dim=(-2, -1), keepdim=True) * rel
    mean = mean * rel
    variance = torch.sum((projections - mean) ** 2, dim=(-2, -1), keepdim=True)
    variance = variance + (large_shape[0] - small_shape[0]) * (large_shape[1] - small_shape[1]) * mean**2
    return (projections * large_size_sqrt) / torch.sqrt(variance.clamp_min(1e-8))


def make_projections(template_dft: torch.Tensor, rotation_matrices: torch.Tensor,
                     projective_filters: torch.Tensor, image_shape, mag_matrix=None) -> torch.Tensor:
    """Slice -> [magnification] -> ifftshift(y) -> DC=0 -> flip -> filter -> irfft2 ->
    ifftshift -> normalise.

    ``backend/cross_correlation.py:330-359``; returns (C, D, B, h, w).
    """
    h = template_dft.shape[1]
    w = 2 * (template_dft.shape[2] - 1)
    s = extract_central_slices_rfft_3d(template_dft, rotation_matrices)
    if mag_matrix is not None:
        s = transform_slice_2d(s, mag_matrix)
    s = torch.fft.ifftshift(s, dim=(-2,))
    s[..., 0, 0] = 0
    s = s * -1
    s = s[None, None, ...] * projective_filters[:, :, None, ...]
    p = torch.fft.irfftn(s, dim=(-2, -1))
    p = torch.fft.ifftshift(p, dim=(-2, -1))
    return normalize_template_projection(p, (h, w), tuple(image_shape))


def batched_cross_correlate(image_dft: torch.Tensor, template_dft: torch.Tensor,
                            rotation_matrices: torch.Tensor, projective_filters: torch.Tensor,
                            mag_matrix=None) -> torch.Tensor:
    """``backend/cross_correlation.py:287-369`` -> (C, D, B, H, W) correlation maps."""
    image_shape = (image_dft.shape[0], 2 * (image_dft.shape[1] - 1))
    p = make_projections(template_dft, rotation_matrices, projective_filters, image_shape, mag_matrix)
    pd = torch.fft.rfftn(p, dim=(-2, -1), s=image_shape)
    pd[..., 0, 0] = 0
    pd = image_dft[None, None, None, ...] * pd.conj()
    return torch.fft.irfftn(pd, dim=(-2, -1), s=image_shape)


def do_iteration_statistics_updates(cc, current_indexes, mip, best_global_index, corr_sum, corr_sq_sum):
    """``backend/utils.py:191-256`` (in place)."""
    h, w = mip.shape
    ccr = cc.reshape(-1, h, w)
    max_values, max_indices = torch.max(ccr, dim=0)
    mask = max_values > mip
    mip.copy_(torch.where(mask, max_values, mip))
    best_global_index.copy_(torch.where(mask, current_indexes[max_indices], best_global_index))
    corr_sum += ccr.sum(dim=0)
    corr_sq_sum += (ccr**2).sum(dim=0)


def aggregate_distributed_results(results: list[dict]) -> dict:
    """``backend/process_results.py:7-64``: max (first device on ties) + sums."""
    mips = np.stack([np.asarray(r["mip"]) for r in results], 0)
    idx = np.stack([np.asarray(r["best_global_index"]) for r in results], 0)
    am = mips.argmax(axis=0)
    return {
        "mip": torch.from_numpy(mips.max(axis=0)),
        "best_global_index": torch.from_numpy(np.take_along_axis(idx, am[None], 0)[0]),
        "correlation_sum": torch.from_numpy(np.stack([np.asarray(r["correlation_sum"]) for r in results]).sum(0)),
        "correlation_squared_sum": torch.from_numpy(
            np.stack([np.asarray(r["correlation_squared_sum"]) for r in results]).sum(0)),
    }


def decode_global_search_index(global_indices, pixel_values, defocus_values, euler_angles):
    """``backend/process_results.py:67-95``."""
    nd = defocus_values.shape[0]
    no = euler_angles.shape[0]
    rem = global_indices.long() % (nd * no)
    d = rem // no
    o = rem % no
    return euler_angles[o, 0], euler_angles[o, 1], euler_angles[o, 2], defocus_values[d]


def scale_mip(mip, corr_sum, corr_sq_sum, total):
    """``backend/process_results.py:98-183``: returns (mip, scaled, mean, std)."""
    mean = corr_sum / total
    var = corr_sq_sum / total
    var = var - mean**2
    std = torch.sqrt(torch.clamp(var, min=0))
    scaled = mip - mean
    scaled = torch.where(std != 0, scaled / std, torch.zeros_like(scaled))
    return mip, scaled, mean, std


def core_match_template(image_dft, template_dft, ctf_filters, whitening_filter_template,
                        defocus_values, pixel_values, euler_angles, orientation_batch_size=1,
                        dtype=torch.float32, return_raw=False, mag_matrix=None):
    """Single-device CPU restatement of ``backend/core_match_template.py:146-363``.

    The orientation loop follows ``:492-554`` with one device, so the work queue is
    the contiguous range [0, O).  Global indices use the (C, D, B) flattening of
    ``:458-466``/``:513-520``.  The final partial batch uses its true size (the
    reference's arange(i, i+B) at ``:513-518`` over-runs it when the last batch is
    short); the two agree whenever O % B == 0 or C*D == 1.
    """
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    image_dft = image_dft.to(cdt)
    template_dft = template_dft.to(cdt)
    filters = (ctf_filters.to(dtype) * whitening_filter_template.to(dtype)[None, None])
    H = image_dft.shape[0]
    W = 2 * (image_dft.shape[1] - 1)
    O = euler_angles.shape[0]
    D = defocus_values.shape[0]
    C = pixel_values.shape[0]
    mip = torch.full((H, W), -float("inf"), dtype=dtype)
    best = torch.full((H, W), -1, dtype=torch.int32)
    csum = torch.zeros((H, W), dtype=dtype)
    csq = torch.zeros((H, W), dtype=dtype)
    incr = torch.tensor([d * O + c * D * O for c in range(C) for d in range(D)], dtype=torch.int32)
    for i in range(0, O, orientation_batch_size):
        eb = euler_angles[i:i + orientation_batch_size]
        rot = euler_to_rotmat_zyz(eb, torch.float64 if dtype == torch.float64 else torch.float32)
        idx = torch.arange(i, i + eb.shape[0], dtype=torch.int32)
        bidx = (idx[None, :] + incr[:, None]).flatten()
        cc = batched_cross_correlate(image_dft, template_dft, rot, filters, mag_matrix)
        do_iteration_statistics_updates(cc, bidx, mip, best, csum, csq)
    if return_raw:
        return {"mip": mip, "best_global_index": best, "correlation_sum": csum, "correlation_squared_sum": csq}
    phi, theta, psi, dfv = decode_global_search_index(best, pixel_values, defocus_values, euler_angles)
    total = O * D * C
    mip, scaled, mean, std = scale_mip(mip, csum, csq, total)
    return {
        "mip": mip, "scaled_mip": scaled, "best_phi": phi, "best_theta": theta, "best_psi": psi,
        "best_defocus": dfv, "correlation_mean": mean, "correlation_variance": std,
        "total_projections": total, "total_orientations": O, "total_defocus": D,
    }


# ------------------------------------------------------------------------- refinement
def refine_particle(particle_image_dft, template_dft, euler_angle, euler_angle_offsets,
                    defocus_u, defocus_v, defocus_angle, defocus_offsets, pixel_size_offsets,
                    corr_mean, corr_std, ctf_kwargs, projective_filter, batch_size=32,
                    dtype=torch.float32, mag_matrix=None):
    """``backend/core_refine_template.py:555-767`` for one particle (CPU branch).

    Returns the dict of per-particle statistics the reference builds at ``:754-767``.
    The batch loop keeps a candidate only when its z-score max is strictly greater
    than the best so far (``:725``); ties inside a batch resolve to the first flat
    index (torch.max on CPU).
    """
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    X = particle_image_dft.to(cdt)
    tdft = template_dft.to(cdt)
    img_h = X.shape[0]
    img_w = 2 * (X.shape[1] - 1)
    th = tdft.shape[1]
    tw = 2 * (tdft.shape[2] - 1)
    crop_h, crop_w = img_h - th + 1, img_w - tw + 1
    best = dict(max_cc=-1e9, max_z_score=-1e9, refined_phi_offset=0.0, refined_theta_offset=0.0,
                refined_psi_offset=0.0, refined_defocus_offset=0.0, refined_pixel_size_offset=0.0,
                refined_pos_y=0, refined_pos_x=0, angle_idx=0)
    rdt = torch.float64 if dtype == torch.float64 else torch.float32
    default_rot = euler_to_rotmat_zyz(euler_angle, rdt)
    kw = {k: v for k, v in ctf_kwargs.items()
          if k in ("voltage", "spherical_aberration", "amplitude_contrast_ratio", "phase_shift", "pixel_size",
                   "ctf_B_factor")}
    ctf = calculate_ctf_filter_stack_full_args(
        (th, tw), float(defocus_u), float(defocus_v), defocus_offsets, pixel_size_offsets,
        astigmatism_angle=float(defocus_angle), mag_matrix=mag_matrix, dtype=torch.float64, **kw).to(dtype)
    filt = projective_filter.to(dtype)[None, None] * ctf
    nb = math.ceil(euler_angle_offsets.shape[0] / batch_size)
    for i in range(nb):
        s0 = i * batch_size
        e0 = min((i + 1) * batch_size, euler_angle_offsets.shape[0])
        ob = euler_angle_offsets[s0:e0]
        rot = euler_to_rotmat_zyz(ob, rdt) @ default_rot
        cc = batched_cross_correlate(X, tdft, rot, filt, mag_matrix)[..., :crop_h, :crop_w]
        z = (cc - corr_mean.to(dtype)) / corr_std.to(dtype)
        if z.max() > best["max_z_score"]:
            best["max_cc"] = float(cc.max())
            best["max_z_score"] = float(z.max())
            mv, mi = torch.max(z.reshape(-1, crop_h, crop_w), dim=0)
            _, mp = torch.max(mv.reshape(-1), dim=0)
            y, x = int(mp) // crop_w, int(mp) % crop_w
            f = int(mi[y, x])
            nbatch = ob.shape[0]
            px = f // (len(defocus_offsets) * nbatch)
            di = (f // nbatch) % len(defocus_offsets)
            ai = f % nbatch
            best.update(
                refined_phi_offset=float(ob[ai, 0]), refined_theta_offset=float(ob[ai, 1]),
                refined_psi_offset=float(ob[ai, 2]), refined_defocus_offset=float(defocus_offsets[di]),
                refined_pixel_size_offset=float(pixel_size_offsets[px]), refined_pos_y=y,
                refined_pos_x=x, angle_idx=ai + s0)
    return best


def core_refine_template(particle_stack_dft, template_dft, euler_angles, euler_angle_offsets,
                         defocus_offsets, defocus_u, defocus_v, defocus_angle, pixel_size_offsets,
                         corr_mean, corr_std, ctf_kwargs, projective_filters, batch_size=32,
                         dtype=torch.float32, mag_matrix=None):
    """``backend/core_refine_template.py:33-202`` on one device, CPU branch; ``mag_matrix``
    enters the slices and the per-particle CTF (``:856-868``, ``:650-660``)."""
    n = particle_stack_dft.shape[0]
    img_h = particle_stack_dft.shape[1]
    img_w = 2 * (particle_stack_dft.shape[2] - 1)
    th = template_dft.shape[1]
    tw = 2 * (template_dft.shape[2] - 1)
    stats = []
    for i in range(n):
        stats.append(refine_particle(
            particle_stack_dft[i], template_dft, euler_angles[i], euler_angle_offsets,
            defocus_u[i], defocus_v[i], defocus_angle[i], defocus_offsets, pixel_size_offsets,
            corr_mean[i], corr_std[i], ctf_kwargs, projective_filters[i], batch_size, dtype, mag_matrix))
    ang = torch.empty((n, 3), dtype=torch.float32)
    for i, st in enumerate(stats):
        off = torch.tensor([st["refined_phi_offset"], st["refined_theta_offset"], st["refined_psi_offset"]],
                           dtype=torch.float32)
        ang[i] = combine_euler_angles(off, euler_angles[i].to(torch.float32))
    ang[:, 0] = torch.where(ang[:, 0] < 0, ang[:, 0] + 360, ang[:, 0])
    ang[:, 1] = torch.where(ang[:, 1] < 0, ang[:, 1] + 180, ang[:, 1])
    ang[:, 2] = torch.where(ang[:, 2] < 0, ang[:, 2] + 360, ang[:, 2])
    pos_y = torch.tensor([st["refined_pos_y"] for st in stats]) - (img_h - th + 1) // 2
    pos_x = torch.tensor([st["refined_pos_x"] for st in stats]) - (img_w - tw + 1) // 2
    return {
        "refined_cross_correlation": torch.tensor([st["max_cc"] for st in stats]),
        "refined_z_score": torch.tensor([st["max_z_score"] for st in stats]),
        "refined_euler_angles": ang,
        "refined_defocus_offset": torch.tensor([st["refined_defocus_offset"] for st in stats]),
        "refined_pixel_size_offset": torch.tensor([st["refined_pixel_size_offset"] for st in stats]),
        "refined_pos_y": pos_y,
        "refined_pos_x": pos_x,
        "angle_idx": torch.tensor([st["angle_idx"] for st in stats]),
    }
