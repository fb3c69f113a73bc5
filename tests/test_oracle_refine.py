"""Pin of the oracle's refinement loop against a brute force (CPU).

``oracle.search.refine_particle`` restates ``backend/core_refine_template.py:555-767``:
batches of orientation offsets composed with the particle's pose, a (pixel size,
defocus, offset) stack of correlations per batch, z = (cc - mean) / std on the valid
crop, the first strict maximum over batches (``:725``), the per-pixel first maximum over
the flattened correlation axis and then the first pixel (``:731-735``), the decode
px = f div (D nb), d = (f div nb) mod D, a = f mod nb (``:738-742``) and max_cc taken as
the batch's whole cc maximum (``:726``).  ``core_refine_template`` adds the angle
composition / wrap (``:504-536``) and the position offset (H - h + 1) // 2 (``:149-191``).

The brute force below computes every correlation on its own (one rotation, one
filter per ``batched_cross_correlate`` call), walks the correlations and pixels in
plain Python loops and applies the rules above literally.  A tie is constructed by
repeating the best offset later in the offset list: the first occurrence must win, both
inside one batch (torch.max keeps the first index) and across batches (strict ``>``).
"""

import math

import pytest
import torch

from oracle import filters as of
from oracle import geometry as og
from oracle import search as osr

CTF_KW = dict(voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0,
              pixel_size=1.0, ctf_B_factor=0.0)


def _inputs(seed, box=32, h=8):
    g = torch.Generator().manual_seed(seed)
    vol = torch.randn((h, h, h), generator=g, dtype=torch.float64)
    tdft = of.volume_to_rfft_fourier_slice(vol)
    img = torch.randn((box, box), generator=g, dtype=torch.float64)
    # plant the template's projection at a known place so that one pose stands out
    X = torch.fft.rfft2(img)
    X[0, 0] = 0
    crop = box - h + 1
    mean = 0.05 * torch.randn((crop, crop), generator=g, dtype=torch.float64)
    std = 1.0 + 0.2 * torch.rand((crop, crop), generator=g, dtype=torch.float64)
    filt = 0.5 + torch.rand((h, h // 2 + 1), generator=g, dtype=torch.float64)
    return X, tdft, mean, std, filt


def _single_cc(X, tdft, rot, filt2d):
    """One correlation map: one rotation (3, 3), one filter (h, h/2+1)."""
    return osr.batched_cross_correlate(X, tdft, rot[None], filt2d[None, None])[0, 0, 0]


def _brute_force(X, tdft, pose, offsets, dfo, pxo, mean, std, filt, batch, du, dv, da):
    h = tdft.shape[1]
    box = X.shape[0]
    crop = box - h + 1
    ctf = of.calculate_ctf_filter_stack_full_args((h, h), du, dv, dfo, pxo, astigmatism_angle=da,
                                                  dtype=torch.float64, **CTF_KW)
    R0 = og.euler_to_rotmat_zyz(pose, torch.float64)
    D, C = len(dfo), len(pxo)
    best = dict(max_cc=-1e9, max_z_score=-1e9, angle_idx=0, pos=(0, 0), d=0, px=0)
    n = offsets.shape[0]
    for s0 in range(0, n, batch):
        nb = min(batch, n - s0)
        # correlations of this batch in the reference's flat (pixel size, defocus, offset) order
        ccs, zs = [], []
        for px in range(C):
            for d in range(D):
                for a in range(nb):
                    rot = og.euler_to_rotmat_zyz(offsets[s0 + a], torch.float64) @ R0
                    cc = _single_cc(X, tdft, rot, filt * ctf[px, d])[:crop, :crop]
                    ccs.append(cc)
                    zs.append((cc - mean) / std)
        zmax = max(float(z.max()) for z in zs)
        if not zmax > best["max_z_score"]:
            continue
        # per pixel: the first correlation reaching the pixel's maximum; then the first pixel
        # (row-major) holding the batch maximum
        pick = None
        for y in range(crop):
            for x in range(crop):
                vals = [float(z[y, x]) for z in zs]
                m = max(vals)
                f = vals.index(m)
                if pick is None or m > pick[0]:
                    pick = (m, y, x, f)
        _, y, x, f = pick
        best.update(max_cc=max(float(c.max()) for c in ccs), max_z_score=zmax, pos=(y, x),
                    px=f // (D * nb), d=(f // nb) % D, angle_idx=s0 + f % nb)
    return best


@pytest.mark.parametrize("batch", [1, 4, 7])
def test_refine_particle_matches_brute_force(batch):
    X, tdft, mean, std, filt = _inputs(3)
    gen = torch.Generator().manual_seed(4)
    offsets = (torch.rand((10, 3), generator=gen, dtype=torch.float64) - 0.5) * 40.0
    pose = torch.tensor([30.0, 70.0, 120.0], dtype=torch.float64)
    dfo = torch.tensor([-300.0, 0.0, 300.0], dtype=torch.float64)
    pxo = torch.tensor([-0.05, 0.05], dtype=torch.float64)
    got = osr.refine_particle(X, tdft, pose, offsets, 12200.0, 11800.0, 30.0, dfo, pxo, mean, std, CTF_KW, filt,
                              batch_size=batch, dtype=torch.float64)
    ref = _brute_force(X, tdft, pose, offsets, dfo, pxo, mean, std, filt, batch, 12200.0, 11800.0, 30.0)
    assert got["angle_idx"] == ref["angle_idx"]
    assert (got["refined_pos_y"], got["refined_pos_x"]) == ref["pos"]
    assert got["refined_defocus_offset"] == float(dfo[ref["d"]])
    assert got["refined_pixel_size_offset"] == float(pxo[ref["px"]])
    assert math.isclose(got["max_z_score"], ref["max_z_score"], rel_tol=1e-12, abs_tol=1e-12)
    assert math.isclose(got["max_cc"], ref["max_cc"], rel_tol=1e-12, abs_tol=1e-12)
    a = ref["angle_idx"]
    assert (got["refined_phi_offset"], got["refined_theta_offset"], got["refined_psi_offset"]) == tuple(
        float(v) for v in offsets[a])


@pytest.mark.parametrize("batch", [1, 4, 7])
def test_refine_particle_ties_keep_first(batch):
    """The best offset repeated later in the list: its first occurrence is reported."""
    X, tdft, mean, std, filt = _inputs(5)
    gen = torch.Generator().manual_seed(6)
    base = (torch.rand((6, 3), generator=gen, dtype=torch.float64) - 0.5) * 40.0
    pose = torch.tensor([10.0, 50.0, 200.0], dtype=torch.float64)
    dfo = torch.tensor([0.0, 250.0], dtype=torch.float64)
    pxo = torch.tensor([0.0], dtype=torch.float64)
    args = (12200.0, 11800.0, 30.0, dfo, pxo, mean, std, CTF_KW, filt)
    first = osr.refine_particle(X, tdft, pose, base, *args, batch_size=len(base), dtype=torch.float64)
    win = first["angle_idx"]
    # the winner again at positions 6 and 8 (same batch as the original for batch 7, later
    # batches for 1 and 4)
    offsets = torch.cat([base, base[win:win + 1], base[:1], base[win:win + 1]])
    got = osr.refine_particle(X, tdft, pose, offsets, *args, batch_size=batch, dtype=torch.float64)
    ref = _brute_force(X, tdft, pose, offsets, dfo, pxo, mean, std, filt, batch, 12200.0, 11800.0, 30.0)
    assert got["angle_idx"] == ref["angle_idx"] == win
    # batched FFTs of different batch compositions agree to rounding only
    assert math.isclose(got["max_z_score"], first["max_z_score"], rel_tol=1e-12)


def test_core_refine_template_position_offset_and_composition():
    """Stack wrapper: per-particle results, pos - (H - h + 1) // 2, composed + wrapped angles."""
    X, tdft, mean, std, filt = _inputs(8)
    X2, _, mean2, std2, filt2 = _inputs(9)
    poses = torch.tensor([[30.0, 70.0, 120.0], [300.0, 20.0, 10.0]], dtype=torch.float64)
    gen = torch.Generator().manual_seed(10)
    offsets = (torch.rand((5, 3), generator=gen, dtype=torch.float64) - 0.5) * 10.0
    dfo = torch.tensor([0.0], dtype=torch.float64)
    pxo = torch.tensor([0.0], dtype=torch.float64)
    du = torch.tensor([12200.0, 15000.0])
    dv = torch.tensor([11800.0, 14800.0])
    da = torch.tensor([30.0, 60.0])
    out = osr.core_refine_template(torch.stack([X, X2]), tdft, poses, offsets, dfo, du, dv, da, pxo,
                                   torch.stack([mean, mean2]), torch.stack([std, std2]), CTF_KW,
                                   torch.stack([filt, filt2]), batch_size=2, dtype=torch.float64)
    box, h = X.shape[0], tdft.shape[1]
    for i, (Xi, mi, si, fi) in enumerate([(X, mean, std, filt), (X2, mean2, std2, filt2)]):
        ref = _brute_force(Xi, tdft, poses[i], offsets, dfo, pxo, mi, si, fi, 2, float(du[i]), float(dv[i]),
                           float(da[i]))
        assert int(out["angle_idx"][i]) == ref["angle_idx"]
        assert int(out["refined_pos_y"][i]) == ref["pos"][0] - (box - h + 1) // 2
        assert int(out["refined_pos_x"][i]) == ref["pos"][1] - (box - h + 1) // 2
        # composed rotation R(offset) @ R(pose), angles wrapped into [0, 360) x [0, 180] x [0, 360)
        rot = og.euler_to_rotmat_zyz(offsets[ref["angle_idx"]]) @ og.euler_to_rotmat_zyz(poses[i])
        e = out["refined_euler_angles"][i]
        assert 0 <= float(e[0]) < 360 and 0 <= float(e[1]) <= 180 and 0 <= float(e[2]) < 360
        assert torch.allclose(og.euler_to_rotmat_zyz(e.double()), rot, atol=1e-5)
