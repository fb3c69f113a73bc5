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


def _ragged_case(seed=9):
    n, box, h = 4, 128, 32
    stack, tdft, poses, mean, std = _particles(n, box, h, seed=seed)
    rng = np.random.default_rng(seed)
    offs, dfos, pxos = [], [], []
    for i, (no, nd, nc) in enumerate([(8, 1, 2), (27, 3, 1), (12, 2, 3), (33, 5, 1)]):
        offs.append(torch.tensor(rng.uniform(-2.0, 2.0, (no, 3)), dtype=torch.float32))
        dfos.append(torch.linspace(-100.0 * (nd // 2), 100.0 * (nd // 2), nd))
        pxos.append(torch.linspace(-0.01 * (nc // 2), 0.01 * (nc // 2), nc))
    du = torch.tensor([12200.0, 15000.0, 11000.0, 13000.0])
    dv = torch.tensor([11800.0, 14700.0, 10900.0, 12500.0])
    da = torch.tensor([30.0, 10.0, 75.0, 0.0])
    filt = torch.ones((n, h, h // 2 + 1))
    return stack, tdft, poses, offs, dfos, du, dv, da, pxos, mean, std, filt


def test_refine_ragged_grids_match_oracle_per_particle(cuda_device):
    """Uneven per-particle grids (offsets 8..33, defocus 1..5, pixel sizes 1..3) in one
    batched call, batches of 8 (full batches from graph replays, partial last batches):
    each particle equals the oracle refining it with its own grid."""
    from leopard_em_amd.backend import refine_template_ragged

    stack, tdft, poses, offs, dfos, du, dv, da, pxos, mean, std, filt = _ragged_case()
    print("ragged refinement: 4 particles, box 128, h 32, batch 8", flush=True)
    got = refine_template_ragged(stack, tdft, poses, offs, dfos, du, dv, da, pxos, mean, std, CTF_KW, filt,
                                 cuda_device, batch_size=8)
    for i in range(stack.shape[0]):
        ref = osr.core_refine_template(stack[i:i + 1], tdft, poses[i:i + 1], offs[i], dfos[i], du[i:i + 1],
                                       dv[i:i + 1], da[i:i + 1], pxos[i], mean[i:i + 1], std[i:i + 1], CTF_KW,
                                       filt[i:i + 1], batch_size=8)
        assert abs(float(got["refined_z_score"][i]) - float(ref["refined_z_score"][0])) < 1e-4 * max(
            1.0, abs(float(ref["refined_z_score"][0]))), i
        for k in ("angle_idx", "refined_pos_y", "refined_pos_x"):
            assert int(got[k][i]) == int(ref[k][0]), (i, k)
        assert float(got["refined_defocus_offset"][i]) == float(ref["refined_defocus_offset"][0])
        assert abs(float(got["refined_pixel_size_offset"][i]) - float(ref["refined_pixel_size_offset"][0])) < 1e-7


def test_refine_batch_graph_replays_bit_identical(cuda_device):
    """The batched refinement with graph replays equals plain launches bit for bit, and
    the shared-grid stack call equals refining each particle on its own."""
    from leopard_em_amd import _lib
    from leopard_em_amd.backend import MatchPlan, refine_template_ragged

    stack, tdft, poses, offs, dfos, du, dv, da, pxos, mean, std, filt = _ragged_case(seed=13)
    outs = []
    for graph in (1, 0):
        old = MatchPlan.__init__

        def init(self, *a, _old=old, _g=graph, **k):
            _old(self, *a, **k)
            self.set_option("graph", _g)

        MatchPlan.__init__ = init
        try:
            outs.append(refine_template_ragged(stack, tdft, poses, offs, dfos, du, dv, da, pxos, mean, std, CTF_KW,
                                               filt, cuda_device, batch_size=8))
        finally:
            MatchPlan.__init__ = old
    for k in outs[0]:
        assert torch.equal(outs[0][k], outs[1][k]), k
    # shared grid: one stack call against one call per particle
    g = offs[1]
    whole = core_refine_template(stack, tdft, poses, g, dfos[1], du, dv, da, pxos[2], mean, std, CTF_KW, filt,
                                 cuda_device, batch_size=8)
    for i in range(stack.shape[0]):
        one = core_refine_template(stack[i:i + 1], tdft, poses[i:i + 1], g, dfos[1], du[i:i + 1], dv[i:i + 1],
                                   da[i:i + 1], pxos[2], mean[i:i + 1], std[i:i + 1], CTF_KW, filt[i:i + 1],
                                   cuda_device, batch_size=8)
        for k in whole:
            assert torch.equal(whole[k][i], one[k][0]), (i, k)
    assert _lib.load() is not None
