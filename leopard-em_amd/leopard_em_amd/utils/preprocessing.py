"""Once-per-search input preparation (the manager side of the 2DTM boundary).

Counterparts of what ``MatchTemplateManager.make_backend_core_function_kwargs``
(``pydantic_models/managers/match_template_manager.py:140-218``) builds before it calls
``core_match_template``: the pre-processed image rfft, the template-grid whitening
filter, the CTF filter stack and the fftshifted 3D template rfft.  These run once per
micrograph and are not part of the timed search.

On CUDA (HIP) tensors the transforms and filters run as the library's own kernels
(``tdtm_rfft2``, ``tdtm_volume_rfft``, ``tdtm_prepare_image``, ``tdtm_ctf_filter_stack``);
there is no fallback, the call raises if the library is not built.  CPU tensors keep the
host computation the reference does (torch on the CPU).

* ``preprocess_image``            ``utils/image_processing.py:16-75``
* ``volume_to_rfft_fourier_slice`` ``utils/image_processing.py:179-202``
* ``whitening_filter``            ``config/correlation_filters.py:49-108`` (radial PSD^-1/2)
* ``ctf_filter_stack``            ``utils/ctf_utils.py:19-116`` (torch_ctf sign convention,
                                  CTF(0) = +amplitude contrast, B-envelope, optional
                                  anisotropic magnification)
* ``bandpass_filter`` / ``arbitrary_curve_filter``
                                  ``config/correlation_filters.py:164-334`` (radial filters in
                                  cycles / pixel; combined with the whitening at :374-410)
"""

from __future__ import annotations

import ctypes
import math

import torch

from .. import _lib

_H = 6.62607015e-34
_ME = 9.1093837015e-31
_E = 1.602176634e-19
_C = 299792458.0


def electron_wavelength_angstrom(voltage_kv: float) -> float:
    v = voltage_kv * 1e3
    return _H / math.sqrt(2 * _ME * _E * v * (1 + _E * v / (2 * _ME * _C**2))) * 1e10


def _rfft_grid(shape, device, dtype):
    h, w = shape
    ky = torch.fft.fftfreq(h, device=device, dtype=dtype)
    kx = torch.fft.rfftfreq(w, device=device, dtype=dtype)
    return torch.meshgrid(ky, kx, indexing="ij")


def ctf_filter_stack(template_shape, defocus_u, defocus_v, astigmatism_angle, defocus_offsets,
                     pixel_size_offsets, *, voltage=300.0, spherical_aberration=2.7,
                     amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0, ctf_B_factor=0.0,
                     mag_matrix=None, device="cpu") -> torch.Tensor:
    """(num_pixel_sizes, num_defocus, h, w//2+1) float32 CTF x B-envelope stack.

    defocus = offsets + (u+v)/2, astigmatism |u-v|/2 at ``astigmatism_angle`` degrees,
    Cs scaled per pixel size as Cs / (p/p0)^4 (``utils/search_utils.py:27-28``).
    ``mag_matrix`` (2 x 2 or 4 values): the CTF at frequency k = (kx, ky) is evaluated at
    A k (``utils/ctf_utils.py:66-100``); the B-envelope stays at k.
    """
    if torch.device(device).type == "cuda":
        return ctf_filter_stack_device(template_shape, defocus_u, defocus_v, astigmatism_angle, defocus_offsets,
                                       pixel_size_offsets, voltage=voltage, spherical_aberration=spherical_aberration,
                                       amplitude_contrast_ratio=amplitude_contrast_ratio, phase_shift=phase_shift,
                                       pixel_size=pixel_size, ctf_B_factor=ctf_B_factor, mag_matrix=mag_matrix,
                                       device=device)
    dt = torch.float64
    dfs = torch.as_tensor(defocus_offsets, dtype=dt, device=device) + 0.5 * (defocus_u + defocus_v)
    astig = abs(defocus_u - defocus_v) / 2
    px_off = torch.as_tensor(pixel_size_offsets, dtype=dt, device=device)
    cs_vals = spherical_aberration / ((pixel_size + px_off) / pixel_size) ** 4
    lam = electron_wavelength_angstrom(voltage)
    ky, kx = _rfft_grid(template_shape, device, dt)
    ky, kx = ky / pixel_size, kx / pixel_size
    env = torch.exp(-ctf_B_factor * (ky**2 + kx**2) / 4)
    if mag_matrix is not None:
        a = torch.as_tensor(mag_matrix, dtype=dt).reshape(2, 2).tolist()
        kx, ky = a[0][0] * kx + a[0][1] * ky, a[1][0] * kx + a[1][1] * ky
    k2 = ky**2 + kx**2
    ang = torch.atan2(ky, kx)
    w = math.atan(amplitude_contrast_ratio / math.sqrt(1 - amplitude_contrast_ratio**2))
    ph = math.radians(phase_shift)
    planes = []
    for cs in cs_vals.tolist():
        df = dfs[:, None, None] + astig * torch.cos(2 * (ang - math.radians(astigmatism_angle)))[None]
        chi = -math.pi * lam * k2[None] * df + math.pi / 2 * cs * 1e7 * lam**3 * k2[None] ** 2 - w - ph
        planes.append(-torch.sin(chi) * env[None])
    return torch.stack(planes, 0).to(torch.float32)


def whitening_filter(image_dft: torch.Tensor, output_shape_rfft, num_freq_bins=None, max_freq=0.5) -> torch.Tensor:
    """1/sqrt(radially averaged |X|^2), max-normalised, resampled to an rfft grid (float32)."""
    dt = torch.float64
    dev = image_dft.device
    hh, wr = image_dft.shape
    ww = 2 * (wr - 1)
    nb = num_freq_bins or (min(hh, ww) // 2 + 1)
    ky, kx = _rfft_grid((hh, ww), dev, dt)
    fr = torch.sqrt(ky**2 + kx**2)
    power = image_dft.abs().to(dt) ** 2
    idx = torch.clamp((fr / max_freq * (nb - 1)).round().long(), 0, nb - 1)
    m = fr <= max_freq
    sums = torch.zeros(nb, dtype=dt, device=dev).index_add_(0, idx[m], power[m])
    cnt = torch.zeros(nb, dtype=dt, device=dev).index_add_(0, idx[m], torch.ones_like(power[m]))
    avg = sums / cnt.clamp_min(1)
    avg = torch.where(avg > 0, avg, torch.full_like(avg, float("inf")))
    curve = avg.rsqrt()
    curve = curve / curve.max()
    h, w2 = output_shape_rfft
    oy, ox = _rfft_grid((h, 2 * (w2 - 1)), dev, dt)
    of = torch.sqrt(oy**2 + ox**2)
    pos = of / max_freq * (nb - 1)
    i0 = torch.clamp(pos.floor().long(), 0, nb - 1)
    i1 = torch.clamp(i0 + 1, 0, nb - 1)
    t = pos - i0.to(dt)
    out = curve[i0] * (1 - t) + curve[i1] * t
    return torch.where(of <= max_freq, out, torch.ones_like(out)).to(torch.float32)


def _radial(shape_real, device=None, dt=torch.float64):
    ky, kx = _rfft_grid(shape_real, device, dt)
    return torch.sqrt(ky**2 + kx**2)


def bandpass_filter(low: float, high: float, falloff: float, shape_real, device=None) -> torch.Tensor:
    """Bandpass on the unshifted rfft grid of an (h, w) real image, float32 (h, w//2+1)
    (``config/correlation_filters.py:236-271``): 1 on low <= |k| <= high (cycles / pixel), a
    raised-cosine edge 0.5 (1 + cos(pi d / falloff)) at distance 0 < d <= falloff outside
    the band, 0 beyond; hard edges for falloff = 0."""
    f = _radial(shape_real, device)
    d = torch.clamp(low - f, min=0) + torch.clamp(f - high, min=0)
    if falloff > 0:
        edge = 0.5 * (1 + torch.cos(math.pi * torch.clamp(d / falloff, max=1.0)))
        out = torch.where(d <= falloff, edge, torch.zeros_like(f))
    else:
        out = (d == 0).to(torch.float64)
    return out.to(torch.float32)


def arbitrary_curve_filter(frequencies, amplitudes, shape_real, device=None) -> torch.Tensor:
    """Radial curve on the unshifted rfft grid, float32 (h, w//2+1)
    (``config/correlation_filters.py:281-334``): linear interpolation of the (frequency,
    amplitude) points at |k| (cycles / pixel), 1 below the first and above the last
    frequency."""
    dt = torch.float64
    fq = torch.as_tensor(frequencies, dtype=dt, device=device).reshape(-1)
    am = torch.as_tensor(amplitudes, dtype=dt, device=device).reshape(-1)
    if fq.numel() < 1 or fq.numel() != am.numel():
        raise ValueError("frequencies and amplitudes must be non-empty and of one length")
    if fq.numel() > 1 and not bool(torch.all(fq[1:] > fq[:-1])):
        raise ValueError("curve frequencies must increase")
    f = _radial(shape_real, device)
    if fq.numel() == 1:
        val = torch.full_like(f, float(am[0]))
    else:
        j = torch.clamp(torch.searchsorted(fq, f, right=True), 1, fq.numel() - 1)
        t = (f - fq[j - 1]) / (fq[j] - fq[j - 1])
        val = am[j - 1] + t * (am[j] - am[j - 1])
    inside = (f >= fq[0]) & (f <= fq[-1])
    return torch.where(inside, val, torch.ones_like(f)).to(torch.float32)


def preprocess_image(image_rfft, cumulative_fourier_filters, bandpass_filter, full_image_shape,
                     extracted_box_shape) -> torch.Tensor:
    """Filter, normalise to unit real-space variance, scale by sqrt(dimensionality)."""
    x = image_rfft * cumulative_fourier_filters
    sq = torch.abs(x) ** 2
    ssum = sq.sum(dim=(-2, -1), keepdim=True) + sq[..., :, 1:-1].sum(dim=(-2, -1), keepdim=True)
    x = x / torch.sqrt(ssum)
    x = x * (bandpass_filter.sum() + bandpass_filter[:, 1:-1].sum()) ** 0.5
    ih, iw = full_image_shape
    bh, bw = extracted_box_shape
    return x * ((ih * iw) / (bh * bw)) ** 0.5


def volume_to_rfft_fourier_slice(volume: torch.Tensor) -> torch.Tensor:
    """fftshift -> rfftn -> fftshift(dims 0, 1) of a cubic volume."""
    if volume.is_cuda:
        return volume_rfft_device(volume)
    v = torch.fft.fftshift(volume, dim=(0, 1, 2))
    v = torch.fft.rfftn(v, dim=(0, 1, 2))
    return torch.fft.fftshift(v, dim=(0, 1))


def prepare_image(image: torch.Tensor, template_size: int, *, whitening: bool = True, num_freq_bins=None,
                  max_freq: float = 0.5, bandpass=None, curve=None):
    """(pre-processed image rfft complex64, template-grid combined filter float32).

    ``match_template_manager.py:160-186``: the image and template filters are whitening x
    bandpass x arbitrary curve on their own grids, the bandpass alone sets the
    normalisation's dimensionality.  ``bandpass`` = (low, high, falloff) or None (off),
    ``curve`` = (frequencies, amplitudes) or None (off).  Defaults: the reference's
    (whitening on, the other filters off).
    """
    if image.is_cuda:
        return prepare_image_device(image, template_size, whitening, num_freq_bins, max_freq, bandpass=bandpass,
                                    curve=curve)
    img_dft = torch.fft.rfftn(image.to(torch.float32))
    img_dft[0, 0] = 0
    shape_img = tuple(image.shape)

    def combined(shape_real, rfft_shape):
        wf = (whitening_filter(img_dft, rfft_shape, num_freq_bins, max_freq) if whitening
              else torch.ones(rfft_shape, dtype=torch.float32))
        if bandpass is not None:
            wf = wf * bandpass_filter(*bandpass, shape_real)
        if curve is not None:
            wf = wf * arbitrary_curve_filter(curve[0], curve[1], shape_real)
        return wf

    wf_img = combined(shape_img, tuple(img_dft.shape))
    wf_tmp = combined((template_size, template_size), (template_size, template_size // 2 + 1))
    bp = (bandpass_filter(*bandpass, shape_img) if bandpass is not None
          else torch.ones(img_dft.shape, dtype=torch.float32))
    pre = preprocess_image(img_dft, wf_img, bp, shape_img, shape_img)
    return pre.to(torch.complex64), wf_tmp


# ---------------------------------------------------------------- device kernels
def _stream(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def rfft2_device(image: torch.Tensor) -> torch.Tensor:
    """Unnormalised rfft2 of a (H, W) float32 CUDA image -> (H, W/2+1) complex64
    (``tdtm_rfft2``; H, W powers of two in [32, 4096])."""
    lib = _lib.load()
    img = image.to(torch.float32).contiguous()
    H, W = img.shape
    out = torch.empty((H, W // 2 + 1), dtype=torch.complex64, device=img.device)
    with torch.cuda.device(img.device):
        _lib.check(lib.tdtm_rfft2(_lib.ptr(img), H, W, _lib.ptr(out), _stream(img.device)))
    return out


def volume_rfft_device(volume: torch.Tensor) -> torch.Tensor:
    """fftshift -> rfftn -> fftshift(dims 0, 1) of a cubic float32 CUDA volume
    (``tdtm_volume_rfft``; side a power of two in [32, 512])."""
    lib = _lib.load()
    vol = volume.to(torch.float32).contiguous()
    n = vol.shape[0]
    if vol.shape != (n, n, n):
        raise ValueError(f"volume must be cubic, got {tuple(vol.shape)}")
    out = torch.empty((n, n, n // 2 + 1), dtype=torch.complex64, device=vol.device)
    with torch.cuda.device(vol.device):
        _lib.check(lib.tdtm_volume_rfft(_lib.ptr(vol), n, _lib.ptr(out), _stream(vol.device)))
    return out


def prepare_image_device(image: torch.Tensor, template_size: int, whitening: bool = True,
                         num_freq_bins: int | None = None, max_freq: float = 0.5, *, bandpass=None, curve=None):
    """(pre-processed image rfft complex64, template-grid combined filter float32) of a CUDA
    image: ``tdtm_prepare_image`` (whitening on or off), or with a ``bandpass`` (low, high,
    falloff) and / or ``curve`` (frequencies, amplitudes) ``tdtm_prepare_image_filtered``."""
    lib = _lib.load()
    img = image.to(torch.float32).contiguous()
    H, W = img.shape
    nb = int(num_freq_bins or 0)
    if bandpass is not None or curve is not None:
        return _prepare_image_filtered(lib, img, template_size, whitening, nb, max_freq, bandpass, curve)
    nbytes = lib.tdtm_prepare_image_workspace_bytes(H, W, nb)
    if nbytes < 0:
        raise _lib.TdtmError(f"unsupported image size {H}x{W} or bin count {nb}")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=img.device)
    out = torch.empty((H, W // 2 + 1), dtype=torch.complex64, device=img.device)
    wf = torch.empty((template_size, template_size // 2 + 1), dtype=torch.float32, device=img.device)
    with torch.cuda.device(img.device):
        _lib.check(lib.tdtm_prepare_image(_lib.ptr(img), H, W, template_size, int(bool(whitening)), nb,
                                          float(max_freq), _lib.ptr(out), _lib.ptr(wf), _lib.ptr(ws),
                                          _stream(img.device)))
    return out, wf


def _prepare_image_filtered(lib, img, template_size, whitening, nb, max_freq, bandpass, curve):
    H, W = img.shape
    nbytes = lib.tdtm_prepare_image_filtered_workspace_bytes(H, W, nb)
    if nbytes < 0:
        raise _lib.TdtmError(f"unsupported image size {H}x{W} or bin count {nb}")
    bp_on, lo, hi, fo = 0, 0.0, 1.0, 0.0
    if bandpass is not None:
        bp_on = 1
        lo, hi, fo = (float(v) for v in bandpass)
    nc, cf, ca = 0, None, None
    if curve is not None:
        fq = [float(v) for v in curve[0]]
        am = [float(v) for v in curve[1]]
        if len(fq) != len(am) or not fq:
            raise ValueError("frequencies and amplitudes must be non-empty and of one length")
        nc = len(fq)
        cf = (ctypes.c_double * nc)(*fq)
        ca = (ctypes.c_double * nc)(*am)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=img.device)
    out = torch.empty((H, W // 2 + 1), dtype=torch.complex64, device=img.device)
    wf = torch.empty((template_size, template_size // 2 + 1), dtype=torch.float32, device=img.device)
    with torch.cuda.device(img.device):
        _lib.check(lib.tdtm_prepare_image_filtered(_lib.ptr(img), H, W, template_size, int(bool(whitening)), nb,
                                                   float(max_freq), bp_on, lo, hi, fo, nc, cf, ca, _lib.ptr(out),
                                                   _lib.ptr(wf), _lib.ptr(ws), _stream(img.device)))
    return out, wf


def ctf_filter_stack_device(template_shape, defocus_u, defocus_v, astigmatism_angle, defocus_offsets,
                            pixel_size_offsets, *, voltage=300.0, spherical_aberration=2.7,
                            amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0, ctf_B_factor=0.0,
                            mag_matrix=None, device="cuda") -> torch.Tensor:
    """``ctf_filter_stack`` on the GPU (``tdtm_ctf_filter_stack`` /
    ``tdtm_ctf_filter_stack_mag``)."""
    lib = _lib.load()
    dev = torch.device(device)
    h = int(template_shape[0])
    if tuple(template_shape) != (h, h):
        raise ValueError("the CTF stack is built for square templates")
    dfo = torch.as_tensor(defocus_offsets, dtype=torch.float32).reshape(-1).to(dev).contiguous()
    pxo = torch.as_tensor(pixel_size_offsets, dtype=torch.float32).reshape(-1).to(dev).contiguous()
    params = (ctypes.c_double * 6)(voltage, spherical_aberration, amplitude_contrast_ratio, phase_shift,
                                        pixel_size, ctf_B_factor)
    out = torch.empty((pxo.numel(), dfo.numel(), h, h // 2 + 1), dtype=torch.float32, device=dev)
    mag = None
    if mag_matrix is not None:
        m = torch.as_tensor(mag_matrix, dtype=torch.float64).cpu().reshape(-1)
        if m.numel() != 4:
            raise ValueError("mag_matrix must have 4 entries (2 x 2)")
        mag = (ctypes.c_double * 4)(*[float(v) for v in m])
    with torch.cuda.device(dev):
        _lib.check(lib.tdtm_ctf_filter_stack_mag(h, dfo.numel(), pxo.numel(), float(defocus_u), float(defocus_v),
                                                 float(astigmatism_angle), _lib.ptr(dfo), _lib.ptr(pxo), params,
                                                 mag, _lib.ptr(out), _stream(dev)))
    return out
