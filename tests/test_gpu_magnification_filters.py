"""GPU parity of anisotropic magnification and of the bandpass / arbitrary-curve
pre-processing filters against the CPU oracle.

* Magnification: ``cross_correlation.py:81-88`` (slices resampled at A k before the
  filters) and the CTF's ``transform_matrix`` (``utils/ctf_utils.py:66-100``), in the search
  and in refinement.  Same tolerances as the plain search (MIP within 1e-4 of max |MIP|
  against the float32 and float64 oracles; refinement z / cc within 1e-4, indices equal).
* Filters: ``config/correlation_filters.py:164-334`` combined with the whitening at
  ``:374-410`` and the bandpass dimensionality of ``utils/image_processing.py:60-67``;
  spectrum and template filter within 1e-5 of the float64 oracle chain.
"""

import numpy as np
import pytest
import torch

from leopard_em_amd import core_match_template, core_refine_template, synthetic
from leopard_em_amd.utils import preprocessing as pp
from oracle import filters as of
from oracle import search as osr

pytestmark = pytest.mark.gpu

MAG = [[1.021, 0.013], [-0.017, 0.987]]


def _rel(a, b):
    a = a.detach().cpu().to(torch.complex128 if a.is_complex() else torch.float64)
    b = b.detach().cpu().to(a.dtype)
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("D,C,bfac", [(1, 1, 0.0), (5, 3, 60.0)])
def test_ctf_stack_with_magnification_matches_oracle(cuda_device, D, C, bfac):
    h = 128
    dfs = torch.linspace(-600.0, 600.0, D) if D > 1 else torch.tensor([0.0])
    pix = torch.linspace(-0.03, 0.03, C) if C > 1 else torch.tensor([0.0])
    kw = dict(voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0,
              pixel_size=1.06, ctf_B_factor=bfac)
    mag = torch.tensor(MAG, dtype=torch.float64)  # one set of values for both sides (the phase is steep)
    got = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, pix, mag_matrix=mag, device=cuda_device, **kw)
    want = of.calculate_ctf_filter_stack_full_args((h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0,
                                                   mag_matrix=mag, **kw)
    plain = of.calculate_ctf_filter_stack_full_args((h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, **kw)
    assert float((got.cpu().double() - want).abs().max()) <= 1e-6
    assert float((plain - want).abs().max()) > 1e-2  # the matrix changes the stack
    # the host restatement of the product (CPU tensors) agrees as well
    host = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, pix, mag_matrix=MAG, **kw)
    assert float((host.double() - want).abs().max()) <= 1e-6


def _match_inputs(H, W, h, D, O, seed):
    t = synthetic.make_blob_template(seed, h, 200, 0.35 * h)
    mg = synthetic.make_micrograph(t, seed, (H, W), 3, 12200.0, 11800.0, 30.0, snr=0.5)
    img_dft, wf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    tdft = of.volume_to_rfft_fourier_slice(t.volume())
    dfs = torch.linspace(-200.0, 200.0, D)
    pix = torch.tensor([0.0])
    ctf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0,
        mag_matrix=torch.tensor(MAG, dtype=torch.float64)).float()
    eul = torch.tensor(np.concatenate([mg.eulers, synthetic.random_uniform_rotations(
        np.random.default_rng(seed + 10), O - 3)]), dtype=torch.float32)
    return (img_dft.to(torch.complex64), tdft.to(torch.complex64), ctf, wf.float(), dfs, pix, eul)


@pytest.mark.parametrize("H,W,h,D,O,B", [(256, 256, 64, 3, 12, 4), (1024, 512, 128, 2, 8, 4)])
def test_match_with_magnification_matches_oracle(cuda_device, H, W, h, D, O, B):
    print(f"match with magnification H{H} W{W} h{h} D{D} O{O} B{B}", flush=True)
    args = _match_inputs(H, W, h, D, O, seed=H + h)
    mag = torch.tensor(MAG)
    got = core_match_template(*args, device=cuda_device, orientation_batch_size=B, mag_matrix=mag)
    r32 = osr.core_match_template(*args, orientation_batch_size=B, mag_matrix=mag)
    r64 = osr.core_match_template(*args, orientation_batch_size=B, dtype=torch.float64,
                                  mag_matrix=mag.double())
    scale = float(r64["mip"].abs().max())
    e32 = float((got["mip"].double() - r32["mip"].double()).abs().max()) / scale
    e64 = float((got["mip"].double() - r64["mip"]).abs().max()) / scale
    print(f"mip rel err vs fp32 {e32:.2e} vs fp64 {e64:.2e}", flush=True)
    assert e32 <= 1e-4 and e64 <= 1e-4
    for k in ("correlation_mean", "correlation_variance"):
        err = float((got[k].double() - r64[k]).abs().max() / r64[k].abs().max())
        assert err <= 1e-4, k
    # without the matrix the maps differ well beyond the tolerance: it is applied
    plain = osr.core_match_template(*args, orientation_batch_size=B)
    assert float((plain["mip"].double() - r32["mip"].double()).abs().max()) / scale > 1e-3


def test_refine_with_magnification_matches_oracle(cuda_device):
    from test_gpu_refine import CTF_KW, _particles

    n, box, h = 3, 128, 32
    stack, tdft, poses, mean, std = _particles(n, box, h, seed=11)
    g = np.arange(-2.0, 2.01, 2.0)
    offs = torch.tensor([[a, b, c] for a in g for b in g for c in g], dtype=torch.float32)
    dfo = torch.tensor([-100.0, 0.0, 100.0])
    pxo = torch.tensor([-0.01, 0.0])
    du = torch.tensor([12200.0] * n)
    dv = torch.tensor([11800.0] * n)
    da = torch.tensor([30.0] * n)
    filt = torch.ones((n, h, h // 2 + 1))
    kw = dict(CTF_KW, mag_matrix=torch.tensor(MAG))
    args = (stack, tdft, poses, offs, dfo, du, dv, da, pxo, mean, std, kw, filt)
    got = core_refine_template(*args, device=cuda_device, batch_size=8, mag_matrix=torch.tensor(MAG))
    ref = osr.core_refine_template(*args, batch_size=8, mag_matrix=torch.tensor(MAG))
    print({k: (got[k][:3], ref[k][:3]) for k in ("refined_z_score", "angle_idx")}, flush=True)
    assert torch.allclose(got["refined_z_score"], ref["refined_z_score"], rtol=1e-4, atol=1e-4)
    assert torch.allclose(got["refined_cross_correlation"], ref["refined_cross_correlation"], rtol=1e-4, atol=1e-4)
    for k in ("angle_idx", "refined_pos_y", "refined_pos_x"):
        assert got[k].tolist() == ref[k].tolist(), k
    assert torch.allclose(got["refined_defocus_offset"], ref["refined_defocus_offset"])
    # the argument and the CTF keyword must agree
    with pytest.raises(ValueError, match="differ"):
        core_refine_template(*args, device=cuda_device, batch_size=8, mag_matrix=torch.eye(2))


@pytest.mark.parametrize("shape,h,bandpass,curve", [
    ((512, 512), 64, (0.02, 0.35, 0.05), None),
    ((256, 512), 64, None, ([0.05, 0.2, 0.45], [1.0, 0.6, 1.4])),
    ((1024, 1024), 128, (0.0, 0.25, 0.0), ([0.1, 0.3], [0.8, 0.4])),
    ((300, 200), 32, (0.05, 0.4, 0.1), ([0.2], [0.5])),
])
def test_prepare_image_filtered_matches_oracle(cuda_device, shape, h, bandpass, curve):
    print(f"filtered preparation {shape} h {h} bandpass {bandpass} curve {curve}", flush=True)
    t = synthetic.make_blob_template(3, h, 80, 0.35 * h)
    mg = synthetic.make_micrograph(t, 5, shape, 4, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32)
    got, wf = pp.prepare_image(img.to(cuda_device), h, bandpass=bandpass, curve=curve)
    torch.cuda.synchronize()
    want, wf_want = of.prepare_image_dft(img.double(), (h, h // 2 + 1), bandpass=bandpass, curve=curve)
    assert _rel(got, want) <= 1e-5
    assert _rel(wf, wf_want) <= 1e-5
    # the bandpass dimensionality: full-spectrum energy sum(bp) + sum(bp[:, 1:-1])
    bp = of.bandpass_filter(*bandpass, shape) if bandpass else torch.ones(got.shape, dtype=torch.float64)
    dim = float(bp.sum() + bp[:, 1:-1].sum())
    p = got.abs().double().cpu() ** 2
    assert abs(float(p.sum() + p[:, 1:-1].sum()) / dim - 1.0) <= 1e-5
    # the host restatement of the product (CPU tensors) agrees with the device
    host, wf_host = pp.prepare_image(img, h, bandpass=bandpass, curve=curve)
    assert _rel(host, want) <= 1e-5 and _rel(wf_host, wf_want) <= 1e-5


def test_prepare_image_filtered_all_pass_is_unfiltered(cuda_device):
    t = synthetic.make_blob_template(4, 64, 60, 20.0)
    mg = synthetic.make_micrograph(t, 6, (512, 512), 3, 12200.0, 11800.0, 30.0, snr=0.3)
    img = mg.image.to(torch.float32).to(cuda_device)
    a, wa = pp.prepare_image_device(img, 64)
    b, wb = pp.prepare_image_device(img, 64, bandpass=(0.0, 1.0, 0.0), curve=([0.0, 1.0], [1.0, 1.0]))
    torch.cuda.synchronize()
    assert _rel(b, a) <= 1e-6 and _rel(wb, wa) <= 1e-7


def test_prepare_image_filtered_refuses_bad_curves(cuda_device):
    from leopard_em_amd import _lib

    img = torch.zeros((64, 64), device=cuda_device)
    with pytest.raises(_lib.TdtmError, match="increase"):
        pp.prepare_image_device(img, 32, curve=([0.2, 0.1], [1.0, 1.0]))
    with pytest.raises(_lib.TdtmError, match="at most"):
        pp.prepare_image_device(img, 32, curve=(list(np.linspace(0, 0.5, 65)), [1.0] * 65))
