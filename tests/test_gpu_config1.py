"""The reference's CPU-runnable plain-search case at its own recipe, on the GPU.

64^3 template of 500 isotropic Gaussians (sigma 1.5 px, amplitude U(0.5, 1), centres
uniform in a radius-24-voxel ball, 1.0 A/px); 512 x 512 micrograph of 20 projections at
uniform SO(3) poses and uniform positions, CTF 300 kV, Cs 2.7 mm, amplitude contrast 0.07,
defocus 12000 A with 200 A astigmatism at 30 deg (defocus_u / v = 12200 / 11800, since the
reference takes astigmatism = |u - v| / 2, utils/ctf_utils.py:51-52), Gaussian noise at SNR
0.1; whitening from the image's own radial power spectrum; the uniform grid at 10 deg
out-of-plane / 5 deg in-plane; one defocus value; orientation batches of 8
(core_match_template.py:146-363 with the manager's inputs, match_template_manager.py:140-218).

* Parity on a strided slice of the grid against the float32 and float64 oracles (MIP,
  mean, std within 1e-4; the z-score peak list identical, with the angles at its peaks).
* The whole grid on the GPU (29,664 orientations): the embedded particles stand out of the
  z-score map although their poses are only approximated by the grid -- at SNR 0.1 a 64^3
  template gives particle z-scores of about 4.5 to 8, so the test checks that the particles
  rank far above the background and that the strongest ones pass the reference's
  one-false-positive cutoff (analysis/zscore_metric.py:14-40), and prints how many pass
  z > 7.5.
"""

import numpy as np
import pytest
import torch

from leopard_em_amd import core_match_template, synthetic
from leopard_em_amd.analysis import find_peaks_from_zscore
from oracle import filters as of
from oracle import geometry as og
from oracle import peaks as op
from oracle import search as osr

pytestmark = pytest.mark.gpu

H = W = 512
h = 64


@pytest.fixture(scope="module")
def config1():
    t = synthetic.make_blob_template(101, h, 500, 24.0)
    mg = synthetic.make_micrograph(t, 101, (H, W), 20, 12200.0, 11800.0, 30.0, snr=0.1)
    img_dft, wf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    tdft = of.volume_to_rfft_fourier_slice(t.volume())
    dfs = torch.tensor([0.0])
    pix = torch.tensor([0.0])
    ctf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0).float()
    grid = og.get_uniform_euler_angles(psi_step=5.0, theta_step=10.0).to(torch.float32)
    args = (img_dft.to(torch.complex64), tdft.to(torch.complex64), ctf, wf.float(), dfs, pix)
    return args, grid, mg


def test_config1_grid_size(config1):
    _, grid, _ = config1
    print(f"10 / 5 degree grid: {grid.shape[0]} orientations", flush=True)
    assert 29000 < grid.shape[0] < 30500


def test_config1_parity_on_grid_slice(config1, cuda_device):
    args, grid, _ = config1
    eul = grid[::32].contiguous()  # 927 orientations spread over the sphere
    got = core_match_template(*args, eul, device=cuda_device, orientation_batch_size=8)
    r32 = osr.core_match_template(*args, eul, orientation_batch_size=8)
    r64 = osr.core_match_template(*args, eul, orientation_batch_size=8, dtype=torch.float64)
    scale = float(r64["mip"].abs().max())
    e32 = float((got["mip"].double() - r32["mip"].double()).abs().max()) / scale
    e64 = float((got["mip"].double() - r64["mip"]).abs().max()) / scale
    print(f"{eul.shape[0]} orientations: mip rel err vs fp32 {e32:.2e} vs fp64 {e64:.2e}", flush=True)
    assert e32 <= 1e-4 and e64 <= 1e-4
    for k in ("correlation_mean", "correlation_variance"):
        err = float((got[k].double() - r64[k]).abs().max() / r64[k].abs().max())
        print(f"{k} rel err vs fp64 {err:.2e}", flush=True)
        assert err <= 1e-4, k
    cut = op.gaussian_noise_zscore_cutoff(H * W * eul.shape[0], 1.0)
    z = r64["scaled_mip"]
    while bool(((z - cut).abs() < 1e-3).any()):  # a cutoff away from every oracle z
        cut += 2e-3
    gy, gx = op.find_peaks_from_zscore(got["scaled_mip"], cut)
    ry, rx = op.find_peaks_from_zscore(z.float(), cut)
    print(f"cutoff {cut:.3f}: {len(ry)} peaks", flush=True)
    assert len(ry) >= 1
    assert gy.tolist() == ry.tolist() and gx.tolist() == rx.tolist()
    for k in ("best_phi", "best_theta", "best_psi"):
        assert got[k][gy, gx].tolist() == r64[k][ry, rx].float().tolist()


def test_config1_full_grid_finds_embedded_particles(config1, cuda_device):
    args, grid, mg = config1
    got = core_match_template(*args, grid, device=cuda_device, orientation_batch_size=8)
    vy, vx = H - h + 1, W - h + 1
    z = got["scaled_mip"][:vy, :vx]
    py, px = find_peaks_from_zscore(z, 7.5)
    peaks = np.stack([py.numpy(), px.numpy()], 1).astype(float) if len(py) else np.zeros((0, 2))
    tl = np.rint(mg.centers_yx).astype(int) - h // 2
    inside = (tl[:, 0] >= 0) & (tl[:, 0] < vy) & (tl[:, 1] >= 0) & (tl[:, 1] < vx)
    found = sum(bool(len(peaks)) and bool((np.hypot(peaks[:, 0] - y, peaks[:, 1] - x) <= 2.0).any())
                for y, x in tl[inside])
    # each particle's z: the best of the 3 x 3 pixels around its template top-left
    zp = np.array([float(z[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2].max()) for y, x in tl[inside]])
    zmap = z.flatten().double()
    q99 = float(torch.quantile(zmap[torch.randperm(zmap.numel(), generator=torch.Generator().manual_seed(0))[:200000]],
                               0.99))
    cut = op.gaussian_noise_zscore_cutoff(H * W * grid.shape[0], 1.0)
    print(f"{grid.shape[0]} orientations: {found} of {int(inside.sum())} particles with a z > 7.5 peak, "
          f"{len(peaks)} peaks; particle z {zp.min():.2f} .. {zp.max():.2f} (median {np.median(zp):.2f}); "
          f"map mean {float(zmap.mean()):.2f}, 99th percentile {q99:.2f}; one-false-positive cutoff {cut:.2f}",
          flush=True)
    assert np.median(zp) > q99
    assert (zp > cut).sum() >= 2
