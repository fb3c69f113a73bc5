"""CTF, envelopes, whitening and pre-processing (CPU oracle, test-only).

Restates:
* ``torch_ctf.calculate_ctf_2d`` and ``torch_fourier_filter.envelopes.b_envelope`` as
  called by ``utils/ctf_utils.py:19-116`` (CTF filter stack).
* ``utils/search_utils.py:6-29`` (Cs for a pixel-size offset).
* ``torch_fourier_filter.whitening.whitening_filter`` as wrapped by
  ``config/correlation_filters.py:49-108``.
* ``utils/image_processing.py:16-75`` (``preprocess_image``) and ``:179-202``
  (``volume_to_rfft_fourier_slice``).
* ``torch_fourier_filter.bandpass.bandpass_filter`` and
  ``torch_fourier_filter.utils.curve_1dim_to_ndim`` as wrapped by
  ``config/correlation_filters.py:164-334`` (bandpass and arbitrary-curve filters).

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
    transform_matrix: torch.Tensor | None = None,
) -> torch.Tensor:
    """CTF on the unshifted rfft grid, one plane per defocus value -> (D, h, w//2+1).

    Astigmatic defocus df(alpha) = df + astig * cos(2 (alpha - alpha_ast)) with
    alpha = atan2(ky, kx) (torch_ctf's convention, defocus_u = df + astig).
    ``transform_matrix`` (anisotropic magnification, ``utils/ctf_utils.py:66-80``,
    ``:97-100``): the CTF at frequency k = (kx, ky) is evaluated at A k, A acting on
    (x, y) column vectors as in :func:`oracle.geometry.transform_slice_2d` (the upstream
    order is parity unpinned).
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
    if transform_matrix is not None:
        a = torch.as_tensor(transform_matrix, dtype=dtype).reshape(2, 2)
        kx, ky = a[0, 0] * kx + a[0, 1] * ky, a[1, 0] * kx + a[1, 1] * ky
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
    mag_matrix: torch.Tensor | None = None,
    dtype=torch.float64,
) -> torch.Tensor:
    """``utils/ctf_utils.py:19-116`` -> (num_pixel_sizes, num_defocus, h, w//2+1).

    defocus = offsets + (u+v)/2, astigmatism = |u-v|/2 (``:51-52``), one Cs per pixel
    size offset via ``get_cs_range`` (``:55-59``), the magnification matrix applied to
    the CTF's frequencies (``:66-100``) but not to the B-envelope (``:103-111``).
    """
    defocus = torch.as_tensor(defocus_offsets, dtype=dtype) + (defocus_u + defocus_v) / 2
    astig = abs(defocus_u - defocus_v) / 2
    cs_values = get_cs_range(pixel_size, torch.as_tensor(pixel_size_offsets, dtype=dtype), spherical_aberration)
    env = b_envelope(ctf_B_factor, template_shape, pixel_size, dtype)
    planes = []
    for cs in cs_values.tolist():
        ctf = calculate_ctf_2d(
            defocus * 1e-4, astig * 1e-4, astigmatism_angle, voltage, cs,
            amplitude_contrast_ratio, phase_shift, pixel_size, template_shape, dtype, mag_matrix,
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


def radial_freq(shape_real: tuple[int, int], dtype=torch.float64) -> torch.Tensor:
    """|k| in cycles / pixel on the unshifted rfft grid of an (h, w) real image
    (torch_grid_utils ``fftfreq_grid(rfft=True, fftshift=False, norm=True)``)."""
    ky, kx = rfft_freq_grid(shape_real, dtype)
    return torch.sqrt(ky**2 + kx**2)


def bandpass_filter(low: float, high: float, falloff: float, shape_real: tuple[int, int],
                    dtype=torch.float64) -> torch.Tensor:
    """Band-pass on the unshifted rfft grid of an (h, w) real image -> (h, w//2+1).

    ``config/correlation_filters.py:264-271`` calls torch_fourier_filter's
    ``bandpass_filter(low, high, falloff, image_shape, rfft=True, fftshift=False)`` with
    cut-offs in cycles / pixel (``from_spatial_resolution``: pixel size / resolution,
    ``:206-212``) and its defaults low 0, high 1, falloff 0 (``:256-259``).  Restated as
    1 on low <= |k| <= high, a raised-cosine edge 0.5 (1 + cos(pi d / falloff)) at a
    distance 0 < d <= falloff outside the band, 0 beyond (hard edges when falloff = 0).
    The upstream edge shape is parity unpinned (the package is absent).
    """
    f = radial_freq(shape_real, dtype)
    d = torch.clamp(low - f, min=0) + torch.clamp(f - high, min=0)  # distance outside the band
    if falloff > 0:
        edge = 0.5 * (1 + torch.cos(math.pi * torch.clamp(d / falloff, max=1.0)))
        return torch.where(d <= falloff, edge, torch.zeros_like(f))
    return (d == 0).to(dtype)


def arbitrary_curve_filter(frequencies, amplitudes, shape_real: tuple[int, int], dtype=torch.float64) -> torch.Tensor:
    """Radial curve on the unshifted rfft grid -> (h, w//2+1),
    ``config/correlation_filters.py:281-334``: ``curve_1dim_to_ndim`` interpolates the
    (frequency, amplitude) pairs linearly at every |k| (cycles / pixel, the same grid as
    :func:`bandpass_filter`), with 1 below the first and above the last frequency
    (``fill_lower`` = ``fill_upper`` = 1).  Frequencies must increase.
    """
    fq = torch.as_tensor(frequencies, dtype=dtype).reshape(-1)
    am = torch.as_tensor(amplitudes, dtype=dtype).reshape(-1)
    if fq.numel() != am.numel() or fq.numel() < 1:
        raise ValueError("frequencies and amplitudes must be non-empty and of one length")
    f = radial_freq(shape_real, dtype)
    j = torch.clamp(torch.searchsorted(fq, f, right=True), 1, max(fq.numel() - 1, 1))
    if fq.numel() == 1:
        val = torch.full_like(f, float(am[0]))
    else:
        f0, f1 = fq[j - 1], fq[j]
        t = (f - f0) / (f1 - f0)
        val = am[j - 1] + t * (am[j] - am[j - 1])
    inside = (f >= fq[0]) & (f <= fq[-1])
    return torch.where(inside, val, torch.ones_like(f))


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


def prepare_image_dft(image: torch.Tensor, template_shape_rfft: tuple[int, int], *, whitening: bool = True,
                      num_freq_bins: int | None = None, max_freq: float = 0.5,
                      bandpass: tuple[float, float, float] | None = None,
                      curve: tuple[list, list] | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Manager-side image preparation, ``match_template_manager.py:160-186``.

    Returns (pre-processed image rfft, template-grid combined filter).  The combined
    filter is whitening x bandpass x arbitrary curve (``correlation_filters.py:374-410``;
    phase randomisation is left out there, ``:395-396``), evaluated on the image grid
    for the image and on the template's rfft grid for the template; the bandpass alone
    sets the dimensionality of ``preprocess_image`` (``:163-165``, ``image_processing.py:60-67``).
    ``bandpass`` = (low, high, falloff) or None (off); ``curve`` = (frequencies,
    amplitudes) or None (off).
    """
    dtype = torch.float64 if image.dtype == torch.float64 else torch.float32
    image_dft = torch.fft.rfftn(image)
    image_dft[0, 0] = 0
    shape_img = tuple(image.shape)
    h, w2 = template_shape_rfft
    shape_tpl = (h, 2 * (w2 - 1))

    def combined(shape_real, rfft_shape):
        if whitening:
            wf = whitening_filter(image_dft, rfft_shape, num_freq_bins, max_freq)
        else:
            wf = torch.ones(rfft_shape, dtype=torch.float64)
        if bandpass is not None:
            wf = wf * bandpass_filter(*bandpass, shape_real)
        if curve is not None:
            wf = wf * arbitrary_curve_filter(curve[0], curve[1], shape_real)
        return wf.to(dtype)

    bp = (bandpass_filter(*bandpass, shape_img) if bandpass is not None
          else torch.ones(image_dft.shape, dtype=torch.float64)).to(dtype)
    wf_image = combined(shape_img, image_dft.shape)
    wf_template = combined(shape_tpl, template_shape_rfft)
    pre = preprocess_image(image_dft, wf_image, bp, shape_img, shape_img)
    return pre, wf_template
