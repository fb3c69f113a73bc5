"""Pins for the oracle's CTF, envelope, whitening and pre-processing (CPU)."""

import math

import numpy as np
import torch

from oracle import filters as of


def test_relativistic_wavelength_textbook_values():
    # 300 kV -> 0.019687 A, 200 kV -> 0.025079 A (standard tables)
    assert abs(of.relativistic_wavelength_angstrom(300.0) - 0.019687) < 2e-6
    assert abs(of.relativistic_wavelength_angstrom(200.0) - 0.025079) < 2e-6


def test_ctf_at_origin_equals_amplitude_contrast():
    ctf = of.calculate_ctf_2d(torch.tensor([1.2]), 0.02, 30.0, 300.0, 2.7, 0.07, 0.0, 1.0, (32, 32))
    assert abs(float(ctf[0, 0, 0]) - 0.07) < 1e-12
    assert float(ctf.abs().max()) <= 1.0 + 1e-12


def test_ctf_zero_rings_without_astigmatism():
    # Cs = 0, A = 0, no astigmatism: CTF = sin(pi lam df k^2) vanishes at k^2 = m/(lam df)
    n, px = 64, 1.0
    lam = of.relativistic_wavelength_angstrom(300.0)
    k_ring = 8 / (n * px)  # eighth rfft sample along x
    df = 1.0 / (lam * k_ring**2)  # first zero exactly on that sample
    ctf = of.calculate_ctf_2d(torch.tensor([df * 1e-4], dtype=torch.float64), 0.0, 0.0, 300.0, 0.0, 0.0, 0.0, px, (n, n))
    assert abs(float(ctf[0, 0, 8])) < 1e-9
    assert abs(float(ctf[0, 8, 0])) < 1e-9
    assert float(ctf[0, 0, 4]) > 0.5  # between zeros: sin(pi/4) ~ 0.707


def test_astigmatism_axis():
    # along the astigmatism angle the defocus is df + astig, perpendicular df - astig
    n = 64
    lam = of.relativistic_wavelength_angstrom(300.0)
    ctf = of.calculate_ctf_2d(torch.tensor([1.0]), 0.1, 0.0, 300.0, 0.0, 0.0, 0.0, 1.0, (n, n))
    k = 5 / n
    for val, dfa in ((float(ctf[0, 0, 5]), 1.1e4), (float(ctf[0, 5, 0]), 0.9e4)):
        assert abs(val - math.sin(math.pi * lam * dfa * k * k)) < 1e-9


def test_b_envelope_closed_form():
    env = of.b_envelope(40.0, (16, 16), 2.0)
    assert float(env[0, 0]) == 1.0
    k = 3 / (16 * 2.0)
    assert abs(float(env[0, 3]) - math.exp(-40.0 * k * k / 4)) < 1e-12
    assert torch.all(of.b_envelope(0.0, (16, 16), 1.0) == 1)


def test_ctf_stack_shapes_and_cs_per_pixel_size():
    st = of.calculate_ctf_filter_stack_full_args(
        (32, 32), 12200.0, 11800.0, torch.tensor([-200.0, 0.0, 200.0]), torch.tensor([-0.02, 0.0]),
        astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07,
        phase_shift=0.0, pixel_size=1.0)
    assert st.shape == (2, 3, 32, 17)
    # different Cs per pixel size changes high frequencies only
    assert float((st[0] - st[1]).abs().max()) > 1e-6


def test_whitening_of_white_noise_is_flat():
    g = torch.Generator().manual_seed(0)
    img = torch.randn((256, 256), generator=g, dtype=torch.float64)
    dft = torch.fft.rfft2(img)
    wf = of.whitening_filter(dft, (64, 33))
    inner = wf[1:20, 1:20]
    assert float(inner.std() / inner.mean()) < 0.15


def test_preprocess_image_unit_variance_parseval():
    # after pre-processing (no bandpass) the full Hermitian spectrum has energy H*W,
    # so by Parseval the real-space image has sum of squares 1
    g = torch.Generator().manual_seed(1)
    img = torch.randn((64, 64), generator=g, dtype=torch.float64) * 3 + 1
    dft = torch.fft.rfft2(img)
    dft[0, 0] = 0
    bp = torch.ones(dft.shape, dtype=torch.float64)
    out = of.preprocess_image(dft, torch.ones_like(bp), bp, (64, 64), (64, 64))
    real = torch.fft.irfft2(out, s=(64, 64))
    assert abs(float((real**2).sum()) - 1.0) < 1e-9


def test_volume_rfft_matches_numpy():
    g = torch.Generator().manual_seed(2)
    vol = torch.randn((8, 8, 8), generator=g, dtype=torch.float64)
    got = of.volume_to_rfft_fourier_slice(vol).numpy()
    ref = np.fft.fftshift(np.fft.rfftn(np.fft.fftshift(vol.numpy())), axes=(0, 1))
    assert np.allclose(got, ref, atol=1e-12)


# ----------------------------------------------- magnification, bandpass, curve filters
def _ctf(ang=30.0, pixel=1.0, mag=None, shape=(32, 32)):
    return of.calculate_ctf_2d(torch.tensor([1.5]), 0.03, ang, 300.0, 2.7, 0.07, 0.0, pixel, shape,
                               transform_matrix=mag)


def test_ctf_identity_magnification_is_no_magnification():
    assert torch.equal(_ctf(mag=torch.eye(2)), _ctf())


def test_ctf_isotropic_magnification_is_a_pixel_size_change():
    # A = s I scales every frequency by s: the CTF at pixel size p with A equals the CTF
    # at pixel size p / s without
    s = 1.02
    got = _ctf(pixel=1.1, mag=s * torch.eye(2, dtype=torch.float64))
    exp = _ctf(pixel=1.1 / s)
    assert float((got - exp).abs().max()) < 1e-12


def test_ctf_rotation_magnification_turns_the_astigmatism_axis():
    # A = R(t) adds t to the polar angle of every frequency and keeps |k|: the same as
    # an astigmatism angle turned by -t
    t = math.radians(17.0)
    r = torch.tensor([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]], dtype=torch.float64)
    got = _ctf(ang=30.0, mag=r)
    exp = _ctf(ang=30.0 - 17.0)
    assert float((got - exp).abs().max()) < 1e-12


def _radial(shape):
    h, w = shape
    ky = np.fft.fftfreq(h)[:, None]
    kx = np.fft.rfftfreq(w)[None, :]
    return np.sqrt(ky**2 + kx**2)


def test_bandpass_band_edges_and_midpoints():
    shape = (64, 64)
    f = _radial(shape)
    lo, hi, fo = 0.1, 0.3, 0.05
    b = of.bandpass_filter(lo, hi, fo, shape).numpy()
    assert b.shape == (64, 33)
    assert np.all(b[(f >= lo) & (f <= hi)] == 1.0)
    assert np.all(b[(f > hi + fo) | (f < lo - fo)] == 0.0)
    assert np.all((b >= 0) & (b <= 1))
    # raised-cosine edges: 0.5 half way, monotone in the distance from the band
    mid = of.bandpass_filter(lo, hi, 2 * abs(f[0, 20] - hi), shape).numpy()[0, 20]
    assert abs(mid - 0.5) < 1e-12
    edge = (f > hi) & (f <= hi + fo)
    order = np.argsort(f[edge])
    assert np.all(np.diff(b[edge][order]) <= 1e-15)


def test_bandpass_limits():
    shape = (32, 48)
    assert torch.all(of.bandpass_filter(0.0, 1.0, 0.0, shape) == 1)  # the disabled defaults pass all
    f = _radial(shape)
    hard = of.bandpass_filter(0.05, 0.2, 0.0, shape).numpy()
    assert np.array_equal(hard, ((f >= 0.05) & (f <= 0.2)).astype(float))
    # the template-grid filter is radial: symmetric under ky -> -ky
    b = of.bandpass_filter(0.05, 0.2, 0.03, (32, 32)).numpy()
    assert np.array_equal(b[1:], b[1:][::-1])


def test_arbitrary_curve_knots_segments_and_fill():
    shape = (64, 64)
    fq, am = [0.1, 0.2, 0.4], [1.0, 0.5, 2.0]
    c = of.arbitrary_curve_filter(fq, am, shape).numpy()
    f = _radial(shape)
    assert np.all(c[f < 0.1] == 1.0) and np.all(c[f > 0.4] == 1.0)
    seg = (f >= 0.2) & (f <= 0.4)
    assert np.allclose(c[seg], 0.5 + (f[seg] - 0.2) / 0.2 * 1.5, atol=1e-12)
    # knots: (0, 6) of a 64-wide grid has |k| = 6/64; (0, 0) row 0 column 8 -> 0.125
    assert abs(c[0, 8] - (1.0 + (0.125 - 0.1) / 0.1 * -0.5)) < 1e-12


def test_preprocess_with_bandpass_has_bandpass_dimensionality():
    # after pre-processing the full Hermitian spectrum's energy equals the bandpass
    # dimensionality sum(bp) + sum(bp[:, 1:-1])
    g = torch.Generator().manual_seed(3)
    img = torch.randn((64, 64), generator=g, dtype=torch.float64)
    pre, wf = of.prepare_image_dft(img, (16, 9), bandpass=(0.05, 0.3, 0.05), curve=([0.1, 0.4], [1.0, 0.3]))
    bp = of.bandpass_filter(0.05, 0.3, 0.05, (64, 64))
    dim = float(bp.sum() + bp[:, 1:-1].sum())
    energy = float((pre.abs() ** 2).sum() + (pre[:, 1:-1].abs() ** 2).sum())
    assert abs(energy / dim - 1) < 1e-10
    # all-pass settings reproduce the default preparation exactly
    p0, w0 = of.prepare_image_dft(img, (16, 9))
    p1, w1 = of.prepare_image_dft(img, (16, 9), bandpass=(0.0, 1.0, 0.0), curve=([0.0, 1.0], [1.0, 1.0]))
    assert torch.allclose(p0, p1, atol=1e-13) and torch.allclose(w0, w1, atol=1e-13)
