"""CTF, envelopes, whitening and pre-processing (CPU oracle, test-only).

Restates:
* ``torch_ctf.calculate_ctf_2d`` and ``torch_fourier_filter.envelopes.b_envelope`` as
  called by ``utils/ctf_utils.py:19-116`` (CTF filter stack).
* ``utils/search_utils.py:6-29`` (Cs for a pixel-size offset).
* ``torch_fourier_filter.whitening.whitening_filter`` as wrapped by
  ``config/correlation_filters.py:49-108``.
* ``utils/image_processing.py:16-75`` (``preprocess_image``) and ``:179-202``
  (``volume_to_rfft_fourier_slice``).

The CTF sign convention follows torch_ctf: CTF(k) = -sin(k1 n4 df + k2 n4^2 - w - phi)
with k1 = -pi*lambda, k2 = pi/2*Cs*lambda^3, w = atan(A/sqrt(1-A^2)), so CTF(0) = +A;
the reference flips template contrast afterwards (``cross_correlation.py:92``).
Parity against the upstream package is unpinned (it is not importable here); the
closed-form values at k = 0 and on the zero rings are pinned by tests.
"""

from __future__ import annotations

import math

import torch

from .geometry import get_cs_range

# CODATA 2018 exact/defined constants
_H = 6.62607015e-34
_ME = 9.1093837015e-31
_E = 1.602176634e-19
_C = 299792458.0


def relativistic_wavelength_angstrom(voltage_kv: float) -> float:
    """Electron wavelength (Angstrom) at an accelerating voltage in kV."""
    v = voltage_kv * 1e3
    lam = _H / math.sqrt(2 * _ME * _E * v * (1 + _E * v / (2 * _ME * _C**2)))
    return lam * 1e10


def rfft_freq_grid(shape: tuple[int, int], dtype=torch.float64) -> tuple[torch.Tensor, torch.Tensor]:
    """(ky, kx) in cycles/pixel for an unshifted rfft of a (h, w) real image."""
    h, w = shape
    ky = torch.fft.fftfreq(h, dtype=dtype)
    kx = torch.fft.rfftfreq(w, dtype=dtype)
    return torch.meshgrid(ky, kx, indexing="ij")


def calculate_ctf_2d(
    defocus_um: torch.Tensor,
    astigmatism_um: float,
    astigmatism_angle_deg: float,
    voltage_kv: float,
    spherical_aberration_mm: float | torch.Tensor,
    amplitude_contrast: float,
    phase_shift_deg: float,
    pixel_size: float,
    image_shape: tuple[int, int],
    dtype=torch.float64,
) -> torch.Tensor:
    """CTF on the unshifted rfft grid, one plane per defocus value -> (D, h, w//2+1).

    Astigmatic defocus df(alpha) = df + astig * cos(2 (alpha - alpha_ast)) with
    alpha = atan2(ky, kx) (torch_ctf's convention, defocus_u = df + astig).
    """
    defocus = torch.as_tensor(defocus_um, dtype=dtype).reshape(-1) * 1e4
    astig = float(astigmatism_um) * 1e4
    ang = math.radians(astigmatism_angle_deg)
    cs = float(spherical_aberration_mm) * 1e7
    lam = relativistic_wavelength_angstrom(voltage_kv)
    k1 = -math.pi * lam
    k2 = math.pi / 2 * cs * lam**3
    k3 = math.atan(amplitude_contrast / math.sqrt(1 - amplitude_contrast**2))
    k5 = math.radians(phase_shift_deg)
    ky, kx = rfft_freq_grid(image_shape, dtype)
    ky = ky / pixel_size
    kx = kx / pixel_size
    n4 = ky**2 + kx**2
    alpha = torch.atan2(ky, kx)
    df = defocus[:, None, None] + astig * torch.cos(2 * (alpha - ang))[None]
    return -torch.sin(k1 * n4[None] * df + k2 * n4[None] ** 2 - k3 - k5)


def b_envelope(b_factor: float, image_shape: tuple[int, int], pixel_size: float, dtype=torch.float64) -> torch.Tensor:
    """exp(-B |k|^2 / 4), |k| in 1/Angstrom, unshifted rfft grid."""
    ky, kx = rfft_freq_grid(image_shape, dtype)
    k2 = (ky**2 + kx**2) / pixel_size**2
    return torch.exp(-b_factor * k2 / 4)


def calculate_ctf_filter_stack_full_args(
    template_shape: tuple[int, int],
    defocus_u: float,
    defocus_v: float,
    defocus_offsets: torch.Tensor,
    pixel_size_offsets: torch.Tensor,
    *,
    astigmatism_angle: float,
    voltage: float,
    spherical_aberration: float,
    amplitude_contrast_ratio: float,
    phase_shift: float,
    pixel_size: float,
    ctf_B_factor: float = 0.0,
    dtype=torch.float64,
) -> torch.Tensor:
    """``utils/ctf_utils.py:19-116`` -> (num_pixel_sizes, num_defocus, h, w//2+1).

    defocus = offsets + (u+v)/2, astigmatism = |u-v|/2 (``:51-52``), one Cs per pixel
    size offset via ``get_cs_range`` (``:55-59``), B-envelope applied (``:103-111``).
    """
    defocus = torch.as_tensor(defocus_offsets, dtype=dtype) + (defocus_u + defocus_v) / 2
    astig = abs(defocus_u - defocus_v) / 2
    cs_values = get_cs_range(pixel_size, torch.as_tensor(pixel_size_offsets, dtype=dtype), spherical_aberration)
    env = b_envelope(ctf_B_factor, template_shape, pixel_size, dtype)
    planes = []
    for cs in cs_values.tolist():
        ctf = calculate_ctf_2d(
            defocus * 1e-4, astig * 1e-4, astigmatism_angle, voltage, cs,
            amplitude_contrast_ratio, phase_shift, pixel_size, template_shape, dtype,
        )
        planes.append(ctf * env[None])
    return torch.stack(planes, 0)


def whitening_filter(
    image_dft: torch.Tensor,
    output_shape_rfft: tuple[int, int],
    num_freq_bins: int | None = None,
    max_freq: float = 0.5,
    dtype=torch.float64,
) -> torch.Tensor:
    """1/sqrt(radially averaged power) of an unshifted rfft, resampled to another grid.

    Bins |f| (cycles/pixel) of the input grid uniformly on [0, max_freq] into
    ``num_freq_bins`` (default ``min(H, W)//2 + 1``) bins, averages |X|^2 per bin,
    takes power^-1/2, normalises the curve to a maximum of 1 and interpolates it
    linearly at |f| of the (h, w//2+1) output grid; |f| > max_freq gets 1.
    Restates ``config/correlation_filters.py:97-108`` -> torch_fourier_filter
    ``whitening_filter``; the upstream binning/normalisation rule is parity unpinned.
    """
    hh, wr = image_dft.shape
    ww = 2 * (wr - 1)
    if num_freq_bins is None:
        num_freq_bins = min(hh, ww) // 2 + 1
    ky, kx = rfft_freq_grid((hh, ww), dtype)
    fr = torch.sqrt(ky**2 + kx**2)
    power = image_dft.abs().to(dtype) ** 2
    idx = torch.clamp((fr / max_freq * (num_freq_bins - 1)).round().long(), 0, num_freq_bins - 1)
    mask = fr <= max_freq
    sums = torch.zeros(num_freq_bins, dtype=dtype).index_add_(0, idx[mask], power[mask])
    cnt = torch.zeros(num_freq_bins, dtype=dtype).index_add_(0, idx[mask], torch.ones_like(power[mask]))
    avg = sums / cnt.clamp_min(1)
    avg = torch.where(avg > 0, avg, torch.full_like(avg, float("inf")))
    curve = avg.rsqrt()
    curve = curve / curve.max()
    h, w2 = output_shape_rfft
    oy, ox = rfft_freq_grid((h, 2 * (w2 - 1)), dtype)
    of = torch.sqrt(oy**2 + ox**2)
    pos = of / max_freq * (num_freq_bins - 1)
    i0 = torch.clamp(pos.floor().long(), 0, num_freq_bins - 1)
    i1 = torch.clamp(i0 + 1, 0, num_freq_bins - 1)
    t = pos - i0.to(dtype)
    out = curve[i0] * (1 - t) + curve[i1] * t
    return torch.where(of <= max_freq, out, torch.ones_like(out))


def preprocess_image(
    image_rfft: torch.Tensor,
    cumulative_fourier_filters: torch.Tensor,
    bandpass_filter: torch.Tensor,
    full_image_shape: tuple[int, int],
    extracted_box_shape: tuple[int, int],
) -> torch.Tensor:
    """``utils/image_processing.py:16-75`` verbatim in arithmetic order."""
    image_rfft = image_rfft * cumulative_fourier_filters
    sq = torch.abs(image_rfft) ** 2
    ssum = torch.sum(sq, dim=(-2, -1), keepdim=True)
    ssum = ssum + torch.sum(sq[..., :, 1:-1], dim=(-2, -1), keepdim=True)
    image_rfft = image_rfft / torch.sqrt(ssum)
    dimensionality = bandpass_filter.sum() + bandpass_filter[:, 1:-1].sum()
    image_rfft = image_rfft * dimensionality**0.5
    img_h, img_w = full_image_shape
    box_h, box_w = extracted_box_shape
    return image_rfft * ((img_h * img_w) / (box_h * box_w)) ** 0.5


def volume_to_rfft_fourier_slice(volume: torch.Tensor) -> torch.Tensor:
    """``utils/image_processing.py:179-202``: fftshift -> rfftn -> fftshift(dims 0, 1)."""
    v = torch.fft.fftshift(volume, dim=(0, 1, 2))
    v = torch.fft.rfftn(v, dim=(0, 1, 2))
    return torch.fft.fftshift(v, dim=(0, 1))


def prepare_image_dft(image: torch.Tensor, template_shape_rfft: tuple[int, int]) -> tuple[torch.Tensor, torch.Tensor]:
    """Manager-side image preparation, ``match_template_manager.py:160-186``.

    Returns (pre-processed image rfft, template-grid whitening filter) with the
    default filter configuration (whitening on, bandpass and curve filters off).
    """
    dtype = torch.float64 if image.dtype == torch.float64 else torch.float32
    image_dft = torch.fft.rfftn(image)
    image_dft[0, 0] = 0
    bandpass = torch.ones(image_dft.shape, dtype=dtype)
    wf_image = whitening_filter(image_dft, image_dft.shape).to(dtype)
    wf_template = whitening_filter(image_dft, template_shape_rfft).to(dtype)
    pre = preprocess_image(image_dft, wf_image, bandpass, tuple(image.shape), tuple(image.shape))
    return pre, wf_template
