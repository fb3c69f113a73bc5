"""GPU parity of the input-preparation kernels (2D / 3D real FFTs, whitening and image
normalisation, CTF stack) against the CPU oracle and numpy's float64 FFT.

Tolerances (float32 device path against float64 references):
* rfft2 / volume rfft: max |difference| <= 2e-6 x max |reference| (float32 FFT
  rounding grows like log2 N; the largest sizes stay below 1e-6).
* pre-processed image spectrum and whitening filters: max |difference| <= 1e-5 x max
  |reference| (the bin assignment is exact; bin means differ by the FFT rounding).
* CTF stack: max |difference| <= 1e-6 (values in [-1, 1], computed in double).
"""

import numpy as np
import pytest
import torch

from leopard_em_amd import _lib, synthetic
from leopard_em_amd.utils import preprocessing as pp
from oracle import filters as of

pytestmark = pytest.mark.gpu


def _rel(a: torch.Tensor, b: torch.Tensor) -> float:
    a = a.detach().cpu().to(torch.complex128 if a.is_complex() else torch.float64)
    b = b.detach().cpu().to(a.dtype)
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("shape", [(32, 32), (64, 128), (256, 64), (1024, 512), (4096, 4096)])
def test_rfft2_matches_numpy(cuda_device, shape):
    print(f"rfft2 {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    img = rng.standard_normal(shape).astype(np.float32)
    got = pp.rfft2_device(torch.from_numpy(img).to(cuda_device))
    torch.cuda.synchronize()
    want = torch.from_numpy(np.fft.rfft2(img.astype(np.float64)))
    assert got.shape == want.shape
    assert _rel(got, want) <= 2e-6


def test_rfft2_impulses(cuda_device):
    """A unit impulse at (y0, x0) has spectrum exp(-2 pi i (ky y0 / H + kx x0 / W))."""
    H, W = 128, 256
    for y0, x0 in [(0, 0), (1, 0), (0, 1), (37, 201)]:
        img = torch.zeros((H, W))
        img[y0, x0] = 1.0
        got = pp.rfft2_device(img.to(cuda_device)).cpu().to(torch.complex128)
        ky = torch.arange(H, dtype=torch.float64)[:, None]
        kx = torch.arange(W // 2 + 1, dtype=torch.float64)[None]
        want = torch.exp(-2j * np.pi * (ky * y0 / H + kx * x0 / W))
        assert float((got - want).abs().max()) <= 2e-6


@pytest.mark.parametrize("n", [32, 64, 128, 256])
def test_volume_rfft_matches_oracle(cuda_device, n):
    print(f"volume rfft {n}", flush=True)
    vol = synthetic.make_blob_template(n, n, 60, 0.35 * n).volume()
    vol = vol + 0.1 * torch.from_numpy(np.random.default_rng(n).standard_normal((n, n, n)).astype(np.float32))
    got = pp.volume_rfft_device(vol.to(cuda_device))
    torch.cuda.synchronize()
    want = of.volume_to_rfft_fourier_slice(vol.double())
    assert got.shape == want.shape
    assert _rel(got, want) <= 2e-6


def test_volume_rfft_512_central_plane(cuda_device):
    """Full-size volume: the kz = 0 plane (index n/2 after the fftshift) is the fftshifted
    2D rfft of the projection along z of the fftshifted volume (Fourier slice theorem,
    exact), and the whole spectrum satisfies Parseval's identity."""
    n = 512
    vol = synthetic.make_blob_template(5, n, 400, 0.35 * n).volume()
    got = pp.volume_rfft_device(vol.to(cuda_device))
    torch.cuda.synchronize()
    v64 = np.fft.fftshift(vol.double().numpy(), axes=(0, 1, 2))
    plane = np.fft.fftshift(np.fft.rfft2(v64.sum(axis=0)), axes=0)
    g = got[n // 2].cpu().to(torch.complex128)
    assert _rel(g, torch.from_numpy(plane)) <= 2e-6
    # Parseval with the Hermitian weights of the half spectrum (kx = 1 .. n/2-1 twice)
    p = got.abs().double() ** 2
    energy = p.sum() + p[..., 1:-1].sum()
    want = float((v64 ** 2).sum()) * n ** 3
    assert abs(float(energy) / want - 1.0) <= 1e-5


@pytest.mark.parametrize("shape,h", [((64, 64), 32), ((256, 512), 64), ((1024, 1024), 128), ((4096, 4096), 512),
                                     ((200, 300), 64), ((201, 300), 32), ((1000, 1500), 128), ((4092, 5760), 512)])
def test_prepare_image_matches_oracle(cuda_device, shape, h):
    print(f"prepare_image {shape} h {h}", flush=True)
    t = synthetic.make_blob_template(3, h, 80, 0.35 * h)
    mg = synthetic.make_micrograph(t, 7, shape, 4, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32)
    got, wf = pp.prepare_image_device(img.to(cuda_device), h)
    torch.cuda.synchronize()
    want, wf_want = of.prepare_image_dft(img.double(), (h, h // 2 + 1))
    assert got.shape == want.shape and wf.shape == wf_want.shape
    assert float(got[0, 0].abs()) == 0.0
    assert _rel(got, want) <= 1e-5
    assert _rel(wf, wf_want) <= 1e-5
    # unit real-space variance (utils/image_processing.py:54-67): sum |X|^2 with the
    # Hermitian weights equals the dimensionality H W
    H, W = shape
    p = got.abs().double().cpu() ** 2
    energy = float(p.sum() + p[:, 1:-1].sum())
    assert abs(energy / (H * W) - 1.0) <= 1e-5


@pytest.mark.parametrize("whitening,nb,fmax", [(True, 100, 0.4), (True, 0, 0.6), (False, 0, 0.5)])
def test_prepare_image_filter_options(cuda_device, whitening, nb, fmax):
    """Whitening off (all-ones filter) and non-default bin counts / cut-off frequencies
    (config/correlation_filters.py:44-47 fields) against the oracle's filter chain."""
    H, W, h = 512, 256, 64
    t = synthetic.make_blob_template(6, h, 60, 20.0)
    mg = synthetic.make_micrograph(t, 8, (H, W), 3, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32)
    got, wf = pp.prepare_image_device(img.to(cuda_device), h, whitening=whitening, num_freq_bins=nb or None,
                                      max_freq=fmax)
    x = torch.fft.rfftn(img.double())
    x[0, 0] = 0
    if whitening:
        f_img = of.whitening_filter(x, tuple(x.shape), nb or None, fmax)
        f_tmp = of.whitening_filter(x, (h, h // 2 + 1), nb or None, fmax)
    else:
        f_img = torch.ones(x.shape, dtype=torch.float64)
        f_tmp = torch.ones((h, h // 2 + 1), dtype=torch.float64)
    want = of.preprocess_image(x, f_img, torch.ones(x.shape, dtype=torch.float64), (H, W), (H, W))
    assert _rel(got, want) <= 1e-5
    assert _rel(wf, f_tmp) <= 1e-5


def test_prepare_image_deterministic(cuda_device):
    t = synthetic.make_blob_template(3, 64, 80, 20.0)
    mg = synthetic.make_micrograph(t, 9, (512, 512), 3, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32).to(cuda_device)
    a, wa = pp.prepare_image_device(img, 64)
    b, wb = pp.prepare_image_device(img, 64)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(wa, wb)


def test_prepare_image_dispatch(cuda_device):
    """prepare_image / volume_to_rfft_fourier_slice / ctf_filter_stack take the device
    kernels for CUDA inputs."""
    t = synthetic.make_blob_template(4, 64, 40, 20.0)
    mg = synthetic.make_micrograph(t, 2, (256, 256), 2, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32).to(cuda_device)
    a, wa = pp.prepare_image(img, 64)
    b, wb = pp.prepare_image_device(img, 64)
    assert a.is_cuda and torch.equal(a, b) and torch.equal(wa, wb)
    v = t.volume().to(cuda_device)
    assert torch.equal(pp.volume_to_rfft_fourier_slice(v), pp.volume_rfft_device(v))


@pytest.mark.parametrize("D,C,bfac", [(1, 1, 0.0), (13, 1, 0.0), (5, 3, 60.0)])
def test_ctf_stack_matches_oracle(cuda_device, D, C, bfac):
    h = 128
    dfs = torch.linspace(-1200.0, 1200.0, D) if D > 1 else torch.tensor([0.0])
    pix = torch.linspace(-0.03, 0.03, C) if C > 1 else torch.tensor([0.0])
    got = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, pix, voltage=300.0, spherical_aberration=2.7,
                              amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.06,
                              ctf_B_factor=bfac, device=cuda_device)
    assert got.is_cuda
    want = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.06, ctf_B_factor=bfac)
    assert got.shape == want.shape
    assert float((got.cpu().double() - want).abs().max()) <= 1e-6


def test_prep_rejects_bad_sizes(cuda_device):
    with pytest.raises(_lib.TdtmError, match="power of two"):
        pp.rfft2_device(torch.zeros((48, 64), device=cuda_device))
    with pytest.raises(_lib.TdtmError, match="power of two"):
        pp.rfft2_device(torch.zeros((64, 16384), device=cuda_device))
    with pytest.raises(_lib.TdtmError, match="power of two"):
        pp.volume_rfft_device(torch.zeros((48, 48, 48), device=cuda_device))
    with pytest.raises(_lib.TdtmError, match="unsupported image size"):
        pp.prepare_image_device(torch.zeros((100, 127), device=cuda_device), 32)
    with pytest.raises(_lib.TdtmError, match="unsupported image size"):
        pp.prepare_image_device(torch.zeros((5000, 128), device=cuda_device), 32)
