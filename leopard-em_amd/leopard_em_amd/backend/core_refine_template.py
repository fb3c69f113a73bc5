"""Particle refinement on the gfx950 HIP path.

Drop-in for ``leopard_em.backend.core_refine_template.core_refine_template``
(``src/leopard_em/backend/core_refine_template.py:33-202``): same signature and
output dictionary.  The whole stack goes to ``tdtm_refine_batch`` in one call: per
particle the CTF stack is built on the device, every batch of ``batch_size`` orientation
offsets goes through the fused slice -> projection -> correlation pipeline (full batches
replayed from captured HIP graphs), and the z-score / argmax epilogue reduces on the
device, so there is no host synchronisation per batch or per particle (the reference
syncs at ``core_refine_template.py:725``).  Only the valid (H-h+1, W-w+1) rows are
transformed along x.  ``refine_template_ragged`` takes a grid of its own per particle.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .core_match_template import MatchPlan, _as_device_list, mag_array

_CTF_KEYS = ("voltage", "spherical_aberration", "amplitude_contrast_ratio", "phase_shift", "pixel_size",
             "ctf_B_factor")


def _ctf_params(ctf_kwargs: dict) -> ctypes.Array:
    for k in ("even_zernikes", "odd_zernikes"):
        z = ctf_kwargs.get(k)
        if z:
            raise NotImplementedError(f"{k} are not supported by the tdtm backend")
    vals = [float(ctf_kwargs.get(k, 0.0)) for k in _CTF_KEYS]
    return (ctypes.c_double * 6)(*vals)


def _magnification(ctf_kwargs: dict, mag_matrix):
    """The one magnification matrix of slices and CTF: the ``mag_matrix`` argument
    (``core_refine_template.py:856-868``) and ``ctf_kwargs["mag_matrix"]`` (``:650-660``),
    which come from the same particle stack in the reference (``utils/backend_setup.py:618-641``)."""
    a, b = mag_array(mag_matrix), mag_array(ctf_kwargs.get("mag_matrix"))
    if a is not None and b is not None and list(a) != list(b):
        raise ValueError("mag_matrix and ctf_kwargs['mag_matrix'] differ")
    return a if a is not None else b


def _int_array(vals) -> ctypes.Array:
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def _double_array(vals) -> ctypes.Array:
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


def _refine_stack(particle_stack_dft, template_dft, euler_angles, offsets, off_start, off_count, dfo, d_start,
                  d_count, pxo, c_start, c_count, defocus_u, defocus_v, defocus_angle, corr_mean, corr_std,
                  ctf_kwargs, projective_filters, device, batch_size, mag_matrix=None) -> dict[str, torch.Tensor]:
    """One tdtm_refine_batch call over the stack; grids given as concatenated arrays with
    per-particle (start, count).  Returns the reference's output dictionary."""
    dev = _as_device_list(device)[0]
    params = _ctf_params(ctf_kwargs)
    mag = _magnification(ctf_kwargs, mag_matrix)
    n, img_h, img_wr = particle_stack_dft.shape
    img_w = 2 * (img_wr - 1)
    th = template_dft.shape[1]
    tw = 2 * (template_dft.shape[2] - 1)
    crop = (img_h - th + 1, img_w - tw + 1)
    if tuple(corr_mean.shape[-2:]) != crop or tuple(corr_std.shape[-2:]) != crop:
        raise ValueError(f"corr_mean / corr_std must have valid-crop shape {crop}")
    D, C = max(d_count), max(c_count)
    plan = MatchPlan(img_h, img_w, th, batch_size, D, C, dev)
    if mag is not None:
        plan.set_magnification(list(mag))
    lib = plan.lib
    with torch.cuda.device(dev):
        st = _lib.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        tdft = template_dft.to(dev, torch.complex64).contiguous()
        if plan.padded:  # crop sizes that are not powers of two: spectra of the padded crops
            stack = torch.stack([plan.regrid(particle_stack_dft[i].to(dev, torch.complex64).contiguous(), st)
                                 for i in range(n)])
        else:
            stack = particle_stack_dft.to(dev, torch.complex64).contiguous()
        stride = stack.shape[1] * stack.shape[2]
        dev_t = [t.to(dev, torch.float32).contiguous() for t in (euler_angles, offsets, dfo, pxo, projective_filters,
                                                                corr_mean, corr_std)]
        poses, offs, dfo_d, pxo_d, filt, mu, sd = dev_t
        out = torch.empty((n, 14), dtype=torch.float32, device=dev)
        _lib.check(lib.tdtm_refine_batch(
            plan.handle, _lib.ptr(tdft), _lib.ptr(stack), n, stride, _lib.ptr(poses), _lib.ptr(offs),
            _int_array(off_start), _int_array(off_count), _lib.ptr(dfo_d), _int_array(d_start), _int_array(d_count),
            _lib.ptr(pxo_d), _int_array(c_start), _int_array(c_count), _double_array(defocus_u),
            _double_array(defocus_v), _double_array(defocus_angle), params, _lib.ptr(filt), _lib.ptr(mu),
            _lib.ptr(sd), _lib.ptr(out), st))
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
    """Refine orientation, defocus, pixel size and position of every particle (one shared
    grid, as the reference; the whole stack in one tdtm_refine_batch call)."""
    _ = num_cuda_streams
    n = particle_stack_dft.shape[0]
    no, nd, nc = euler_angle_offsets.shape[0], defocus_offsets.shape[0], pixel_size_offsets.shape[0]
    return _refine_stack(particle_stack_dft, template_dft, euler_angles, euler_angle_offsets, [0] * n, [no] * n,
                         defocus_offsets, [0] * n, [nd] * n, pixel_size_offsets, [0] * n, [nc] * n,
                         [float(v) for v in defocus_u], [float(v) for v in defocus_v],
                         [float(v) for v in defocus_angle], corr_mean, corr_std, ctf_kwargs, projective_filters,
                         device, batch_size, mag_matrix)


# pylint: disable=too-many-arguments,too-many-locals
def refine_template_ragged(
    particle_stack_dft: torch.Tensor,
    template_dft: torch.Tensor,
    euler_angles: torch.Tensor,
    euler_angle_offsets: list[torch.Tensor],
    defocus_offsets: list[torch.Tensor],
    defocus_u: torch.Tensor,
    defocus_v: torch.Tensor,
    defocus_angle: torch.Tensor,
    pixel_size_offsets: list[torch.Tensor],
    corr_mean: torch.Tensor,
    corr_std: torch.Tensor,
    ctf_kwargs: dict,
    projective_filters: torch.Tensor,
    device: torch.device | list[torch.device],
    batch_size: int = 32,
    mag_matrix: torch.Tensor | None = None,
) -> dict[str, torch.Tensor]:
    """``core_refine_template`` with a grid of its own for every particle: particle i
    searches ``euler_angle_offsets[i]`` (n_i, 3), ``defocus_offsets[i]`` and
    ``pixel_size_offsets[i]`` (any lengths).  No counterpart in the reference, whose grid is
    shared by the stack (``core_refine_template.py:37-42``); per particle the search,
    z-scores, decode and output are those of the reference.  ``angle_idx`` indexes the
    particle's own offset list."""
    n = particle_stack_dft.shape[0]
    if not (len(euler_angle_offsets) == len(defocus_offsets) == len(pixel_size_offsets) == n):
        raise ValueError("one offset, defocus and pixel-size list per particle")

    def concat(ts):
        counts = [int(t.shape[0]) for t in ts]
        starts = [sum(counts[:i]) for i in range(n)]
        return torch.cat([t.reshape(t.shape[0], -1) for t in ts]).squeeze(-1), starts, counts

    offs, o_s, o_c = concat(euler_angle_offsets)
    dfo, d_s, d_c = concat(defocus_offsets)
    pxo, c_s, c_c = concat(pixel_size_offsets)
    return _refine_stack(particle_stack_dft, template_dft, euler_angles, offs.reshape(-1, 3), o_s, o_c, dfo, d_s,
                         d_c, pxo, c_s, c_c, [float(v) for v in defocus_u], [float(v) for v in defocus_v],
                         [float(v) for v in defocus_angle], corr_mean, corr_std, ctf_kwargs, projective_filters,
                         device, batch_size, mag_matrix)
