"""Particle refinement on the gfx950 HIP path.

Drop-in for ``leopard_em.backend.core_refine_template.core_refine_template``
(``src/leopard_em/backend/core_refine_template.py:33-202``): same signature and
output dictionary.  Each particle runs through ``tdtm_refine_particle``: the CTF stack
is built on the device, every batch of ``batch_size`` orientation offsets goes through
the fused slice -> projection -> correlation pipeline, and the z-score / argmax
epilogue reduces on the device, so there is no host synchronisation per batch
(the reference syncs at ``core_refine_template.py:725``).  Only the valid
(H-h+1, W-w+1) rows are transformed along x.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .core_match_template import MatchPlan, _as_device_list

_CTF_KEYS = ("voltage", "spherical_aberration", "amplitude_contrast_ratio", "phase_shift", "pixel_size",
             "ctf_B_factor")


def _ctf_params(ctf_kwargs: dict) -> ctypes.Array:
    for k in ("even_zernikes", "odd_zernikes"):
        z = ctf_kwargs.get(k)
        if z:
            raise NotImplementedError(f"{k} are not supported by the tdtm backend")
    if ctf_kwargs.get("mag_matrix") is not None:
        raise NotImplementedError("anisotropic magnification is not supported by the tdtm backend")
    vals = [float(ctf_kwargs.get(k, 0.0)) for k in _CTF_KEYS]
    return (ctypes.c_double * 6)(*vals)


# pylint: disable=too-many-arguments,too-many-locals
def core_refine_template(
    particle_stack_dft: torch.Tensor,
    template_dft: torch.Tensor,
    euler_angles: torch.Tensor,
    euler_angle_offsets: torch.Tensor,
    defocus_offsets: torch.Tensor,
    defocus_u: torch.Tensor,
    defocus_v: torch.Tensor,
    defocus_angle: torch.Tensor,
    pixel_size_offsets: torch.Tensor,
    corr_mean: torch.Tensor,
    corr_std: torch.Tensor,
    ctf_kwargs: dict,
    projective_filters: torch.Tensor,
    device: torch.device | list[torch.device],
    batch_size: int = 32,
    num_cuda_streams: int = 1,
    mag_matrix: torch.Tensor | None = None,
) -> dict[str, torch.Tensor]:
    """Refine orientation, defocus, pixel size and position of every particle."""
    _ = num_cuda_streams
    if mag_matrix is not None:
        raise NotImplementedError("anisotropic magnification (mag_matrix) is not supported by the tdtm backend")
    dev = _as_device_list(device)[0]
    params = _ctf_params(ctf_kwargs)
    n, img_h, img_wr = particle_stack_dft.shape
    img_w = 2 * (img_wr - 1)
    th = template_dft.shape[1]
    tw = 2 * (template_dft.shape[2] - 1)
    D = defocus_offsets.shape[0]
    C = pixel_size_offsets.shape[0]
    crop = (img_h - th + 1, img_w - tw + 1)
    if tuple(corr_mean.shape[-2:]) != crop or tuple(corr_std.shape[-2:]) != crop:
        raise ValueError(f"corr_mean / corr_std must have valid-crop shape {crop}")
    plan = MatchPlan(img_h, img_w, th, batch_size, D, C, dev)
    lib = plan.lib
    with torch.cuda.device(dev):
        st = _lib.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        tdft = template_dft.to(dev, torch.complex64).contiguous()
        offs = euler_angle_offsets.to(dev, torch.float32).contiguous()
        dfo = defocus_offsets.to(dev, torch.float32).contiguous()
        pxo = pixel_size_offsets.to(dev, torch.float32).contiguous()
        out = torch.empty((n, 14), dtype=torch.float32, device=dev)
        keep = []
        for i in range(n):
            xp = particle_stack_dft[i].to(dev, torch.complex64).contiguous()
            if plan.padded:  # crop sizes that are not powers of two: spectrum of the padded crop
                xp = plan.regrid(xp, st)
            pose = euler_angles[i].to(dev, torch.float32).contiguous()
            filt = projective_filters[i].to(dev, torch.float32).contiguous()
            mu = corr_mean[i].to(dev, torch.float32).contiguous()
            sd = corr_std[i].to(dev, torch.float32).contiguous()
            _lib.check(lib.tdtm_refine_particle(
                plan.handle, _lib.ptr(tdft), _lib.ptr(xp), _lib.ptr(pose), _lib.ptr(offs), offs.shape[0],
                _lib.ptr(dfo), _lib.ptr(pxo), float(defocus_u[i]), float(defocus_v[i]), float(defocus_angle[i]),
                params, _lib.ptr(filt), _lib.ptr(mu), _lib.ptr(sd), _lib.ptr(out[i]), st))
            keep.append((xp, pose, filt, mu, sd))
            if len(keep) > 64:  # bound host-side references; the stream orders reuse
                torch.cuda.current_stream(dev).synchronize()
                keep.clear()
        res = out.cpu()
    pos_y = res[:, 7].round().long() - (img_h - th + 1) // 2
    pos_x = res[:, 8].round().long() - (img_w - tw + 1) // 2
    return {
        "refined_cross_correlation": res[:, 0].clone(),
        "refined_z_score": res[:, 1].clone(),
        "refined_euler_angles": res[:, 2:5].clone(),
        "refined_defocus_offset": res[:, 5].clone(),
        "refined_pixel_size_offset": res[:, 6].clone(),
        "refined_pos_y": pos_y,
        "refined_pos_x": pos_x,
        "angle_idx": res[:, 9].round().long(),
    }
