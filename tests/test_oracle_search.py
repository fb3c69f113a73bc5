"""Pins for the oracle's normalisation, correlation, statistics and result steps (CPU)."""

import math

import numpy as np
import pytest
import torch

from oracle import geometry as og
from oracle import peaks as op
from oracle import search as osr


def test_normalize_is_affine_invariant_and_edge_mean_zero():
    g = torch.Generator().manual_seed(0)
    p = torch.randn((3, 16, 16), generator=g, dtype=torch.float64)
    a = osr.normalize_template_projection(p, (16, 16), (64, 64))
    b = osr.normalize_template_projection(3.5 * p - 7.0, (16, 16), (64, 64))
    assert torch.allclose(a, b, atol=1e-12)
    edge = torch.cat([a[:, 0, :], a[:, -1, :], a[:, 1:-1, 0], a[:, 1:-1, -1]], -1)
    assert torch.allclose(edge.mean(-1), torch.zeros(3, dtype=torch.float64), atol=1e-12)


def test_normalize_padded_variance_large_canvas():
    # for a large canvas the (double-scaled) mean is negligible and the padded
    # canvas has unit second moment: sum(x^2) / (H W) -> 1
    g = torch.Generator().manual_seed(1)
    p = torch.randn((16, 16), generator=g, dtype=torch.float64)
    out = osr.normalize_template_projection(p, (16, 16), (4096, 4096))
    assert abs(float((out**2).sum()) / 4096**2 - 1.0) < 1e-9


def test_normalize_double_relative_size_quirk():
    """backend/utils.py:177-178 scales the mean by rel twice; check its closed-form effect.

    With x the edge-mean-subtracted projection, n = h w, N = H W, rel = n/N and
    m = mean(x): the reference variance is sum((x - m rel^2)^2) + (N - n) ... computed
    with (H-h)(W-w), not the exact padded variance.  Build the padded canvas explicitly
    and check the padded second moment against the prediction.
    """
    g = torch.Generator().manual_seed(2)
    p = torch.randn((8, 8), generator=g, dtype=torch.float64) + 0.3
    H = W = 16
    out = osr.normalize_template_projection(p, (8, 8), (H, W))
    canvas = torch.zeros((H, W), dtype=torch.float64)
    canvas[:8, :8] = out
    x = p - torch.cat([p[0], p[-1], p[1:-1, 0], p[1:-1, -1]]).mean()
    rel = 64 / 256
    mu = float(x.mean()) * rel * rel
    v_ref = float(((x - mu) ** 2).sum()) + (H - 8) * (W - 8) * mu * mu
    scale = math.sqrt(H * W) / math.sqrt(v_ref)
    # canvas second moment of the reference-normalised projection
    assert abs(float((canvas**2).sum()) - scale**2 * float((x**2).sum())) < 1e-9
    assert abs(float(canvas.sum()) - scale * float(x.sum())) < 1e-9


def _direct_circular_correlation(proj_small, img, H, W):
    """cc[n] = sum_m (p_pad[m] - mean(p_pad)) img[(n + m) mod N], float64 brute force."""
    h, w = proj_small.shape
    pad = np.zeros((H, W))
    pad[:h, :w] = proj_small
    pad -= pad.mean()  # P[0,0] = 0 in the reference
    out = np.zeros((H, W))
    for dy in range(H):
        for dx in range(W):
            out[dy, dx] = np.sum(pad * np.roll(np.roll(img, -dy, 0), -dx, 1))
    return out


def test_cross_correlation_matches_direct_correlation():
    H, W, h = 16, 16, 8
    g = torch.Generator().manual_seed(3)
    vol = torch.complex(torch.randn((h, h, h // 2 + 1), generator=g, dtype=torch.float64),
                        torch.randn((h, h, h // 2 + 1), generator=g, dtype=torch.float64))
    img = torch.randn((H, W), generator=g, dtype=torch.float64)
    img_dft = torch.fft.rfft2(img)
    img_dft[0, 0] = 0
    filt = torch.rand((1, 2, h, h // 2 + 1), generator=g, dtype=torch.float64)
    rot = og.euler_to_rotmat_zyz(torch.tensor([[10.0, 20.0, 30.0], [100.0, 80.0, 5.0]]))
    cc = osr.batched_cross_correlate(img_dft, vol, rot, filt)
    proj = osr.make_projections(vol, rot, filt, (H, W))
    img0 = torch.fft.irfft2(img_dft, s=(H, W)).numpy()
    for d in range(2):
        for b in range(2):
            ref = _direct_circular_correlation(proj[0, d, b].numpy(), img0, H, W)
            assert np.allclose(cc[0, d, b].numpy(), ref, atol=1e-10)


def test_projection_quadrant_swap_equals_sign_modulation():
    # ifftshift of the irfft2 output == irfft2 of the spectrum times (-1)^(ky+kx)
    g = torch.Generator().manual_seed(4)
    s = torch.complex(torch.randn((8, 5), generator=g, dtype=torch.float64),
                      torch.randn((8, 5), generator=g, dtype=torch.float64))
    a = torch.fft.ifftshift(torch.fft.irfft2(s), dim=(0, 1))
    ky = torch.arange(8)[:, None]
    kx = torch.arange(5)[None, :]
    b = torch.fft.irfft2(s * ((-1.0) ** (ky + kx)))
    assert torch.allclose(a, b, atol=1e-12)


def test_statistics_update_bruteforce_and_tie_rule():
    rng = np.random.default_rng(5)
    H, W = 4, 5
    mip = torch.full((H, W), -float("inf"))
    best = torch.full((H, W), -1, dtype=torch.int32)
    s = torch.zeros((H, W), dtype=torch.float64)
    q = torch.zeros((H, W), dtype=torch.float64)
    all_cc, all_idx = [], []
    for it in range(3):
        cc = torch.tensor(rng.integers(-3, 4, (1, 2, 3, H, W)).astype(np.float64))  # many ties
        idx = torch.arange(it * 6, it * 6 + 6, dtype=torch.int32)
        osr.do_iteration_statistics_updates(cc, idx, mip, best, s, q)
        all_cc.append(cc.reshape(6, H, W).numpy())
        all_idx.append(idx.numpy())
    stack = np.concatenate(all_cc, 0)
    gidx = np.concatenate(all_idx)
    # first occurrence of the overall max wins (strict > across batches, first in batch)
    exp_best = gidx[np.argmax(stack, axis=0)]
    assert np.array_equal(best.numpy(), exp_best)
    assert np.allclose(mip.numpy(), stack.max(0))
    assert np.allclose(s.numpy(), stack.sum(0))
    assert np.allclose(q.numpy(), (stack**2).sum(0))


def test_decode_global_index_bruteforce():
    O, D, C = 5, 3, 2
    eul = torch.arange(O * 3, dtype=torch.float32).reshape(O, 3)
    dfs = torch.tensor([-100.0, 0.0, 100.0])
    for c in range(C):
        for d in range(D):
            for o in range(O):
                g = torch.tensor([c * D * O + d * O + o])
                phi, theta, psi, df = osr.decode_global_search_index(g, torch.zeros(C), dfs, eul)
                assert float(phi) == float(eul[o, 0]) and float(psi) == float(eul[o, 2])
                assert float(df) == float(dfs[d])


def test_scale_mip_closed_form_and_zero_std():
    mip = torch.tensor([[5.0, 1.0]], dtype=torch.float64)
    s = torch.tensor([[10.0, 4.0]], dtype=torch.float64)
    q = torch.tensor([[30.0, 4.0]], dtype=torch.float64)  # second pixel: zero variance
    _, z, mean, std = osr.scale_mip(mip, s, q, 4)
    assert torch.allclose(mean, torch.tensor([[2.5, 1.0]], dtype=torch.float64))
    assert torch.allclose(std, torch.tensor([[math.sqrt(30 / 4 - 2.5**2), 0.0]], dtype=torch.float64))
    assert float(z[0, 0]) == pytest.approx((5 - 2.5) / math.sqrt(1.25))
    assert float(z[0, 1]) == 0.0


def test_aggregate_first_device_wins_ties():
    a = {"mip": np.array([[1.0, 2.0]]), "best_global_index": np.array([[1, 2]], dtype=np.int32),
         "correlation_sum": np.array([[1.0, 1.0]]), "correlation_squared_sum": np.array([[1.0, 1.0]])}
    b = {"mip": np.array([[1.0, 3.0]]), "best_global_index": np.array([[7, 8]], dtype=np.int32),
         "correlation_sum": np.array([[2.0, 2.0]]), "correlation_squared_sum": np.array([[3.0, 3.0]])}
    r = osr.aggregate_distributed_results([a, b])
    assert r["best_global_index"].tolist() == [[1, 8]]
    assert r["mip"].tolist() == [[1.0, 3.0]]
    assert r["correlation_sum"].tolist() == [[3.0, 3.0]]


def test_zscore_cutoff_inverts_gaussian_tail():
    from scipy.special import erfc

    for n in (10**6, 10**9, 2 * 10**10):
        c = op.gaussian_noise_zscore_cutoff(n, 1.0)
        assert abs(n * 0.5 * erfc(c / math.sqrt(2)) - 1.0) < 1e-6


def test_peak_picking_bruteforce():
    z = torch.zeros((20, 20))
    pts = {(3, 3): 9.0, (3, 5): 8.5, (10, 10): 8.0, (14, 3): 10.0, (13, 4): 7.9}
    for (y, x), v in pts.items():
        z[y, x] = v
    py, px = op.find_peaks_from_zscore(z, 7.5, mask_radius=5.0)
    got = list(zip(py.tolist(), px.tolist()))
    # descending: (14,3) 10 -> suppresses (13,4); (3,3) 9 -> suppresses (3,5); (10,10) 8
    assert got == [(14, 3), (3, 3), (10, 10)]
    py, px = op.find_peaks_from_zscore(z, 20.0)
    assert py.numel() == 0
