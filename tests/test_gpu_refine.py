"""GPU parity of the HIP particle refinement against the CPU oracle."""

import numpy as np
import pytest
import torch

from leopard_em_amd import core_refine_template, synthetic
from oracle import filters as of
from oracle import search as osr

pytestmark = pytest.mark.gpu

CTF_KW = dict(voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0,
              pixel_size=1.0, ctf_B_factor=0.0, even_zernikes=None, odd_zernikes=None)


def _particles(n, box, h, seed):
    t = synthetic.make_blob_template(seed, h, 120, 0.35 * h)
    vol = t.volume()
    tdft = of.volume_to_rfft_fourier_slice(vol).to(torch.complex64)
    rng = np.random.default_rng(seed)
    stack, poses = [], []
    for i in range(n):
        mg = synthetic.make_micrograph(t, seed + 17 * i, (box, box), 1, 12200.0, 11800.0, 30.0, snr=1.0,
                                       margin=box // 2 - 1)
        dft = torch.fft.rfft2(mg.image)
        dft[0, 0] = 0
        dft = dft / dft.abs().pow(2).sum().sqrt() * box
        stack.append(dft.to(torch.complex64))
        poses.append(mg.eulers[0] + rng.uniform(-2, 2, 3))
    crop = box - h + 1
    mean = torch.zeros((n, crop, crop)) + 0.01 * torch.randn((n, crop, crop), generator=torch.Generator().manual_seed(seed))
    std = 1.0 + 0.1 * torch.rand((n, crop, crop), generator=torch.Generator().manual_seed(seed + 1))
    return (torch.stack(stack), tdft, torch.tensor(np.array(poses), dtype=torch.float32), mean, std)


@pytest.mark.parametrize("batch,box,h", [(4, 128, 32), (32, 128, 32), (8, 1024, 64), (8, 200, 64), (8, 528, 128)])
def test_refine_parity_against_oracle(batch, box, h, cuda_device):
    n = 3
    stack, tdft, poses, mean, std = _particles(n, box, h, seed=5)
    g = np.arange(-2.0, 2.01, 2.0)
    offs = torch.tensor([[a, b, c] for a in g for b in g for c in g], dtype=torch.float32)  # 27 offsets
    dfo = torch.tensor([-100.0, 0.0, 100.0])
    pxo = torch.tensor([-0.01, 0.0])
    du = torch.tensor([12200.0] * n)
    dv = torch.tensor([11800.0] * n)
    da = torch.tensor([30.0] * n)
    filt = torch.ones((n, h, h // 2 + 1))
    args = (stack, tdft, poses, offs, dfo, du, dv, da, pxo, mean, std, CTF_KW, filt)
    got = core_refine_template(*args, device=cuda_device, batch_size=batch)
    ref = osr.core_refine_template(*args, batch_size=batch)
    print({k: (got[k][:3], ref[k][:3]) for k in ("refined_z_score", "angle_idx", "refined_pos_y")}, flush=True)
    assert torch.allclose(got["refined_z_score"], ref["refined_z_score"], rtol=1e-4, atol=1e-4)
    assert torch.allclose(got["refined_cross_correlation"], ref["refined_cross_correlation"], rtol=1e-4, atol=1e-4)
    for k in ("angle_idx", "refined_pos_y", "refined_pos_x"):
        assert got[k].tolist() == ref[k].tolist(), k
    assert torch.allclose(got["refined_defocus_offset"], ref["refined_defocus_offset"])
    assert torch.allclose(got["refined_pixel_size_offset"], ref["refined_pixel_size_offset"])
    # composed angles: compare rotations (Euler wrap is not unique at gimbal points)
    from oracle import geometry as og
    rg = og.euler_to_rotmat_zyz(got["refined_euler_angles"].double())
    rr = og.euler_to_rotmat_zyz(ref["refined_euler_angles"].double())
    assert torch.allclose(rg, rr, atol=1e-4)
