"""Pins for the oracle's rotation, slice-extraction and grid restatements (CPU)."""

import math

import numpy as np
import pytest
import torch
from scipy import ndimage

from oracle import geometry as og


def test_rotation_orthonormal_det1():
    eul = torch.tensor(np.random.default_rng(0).uniform([0, 0, 0], [360, 180, 360], (50, 3)))
    r = og.euler_to_rotmat_zyz(eul)
    eye = torch.eye(3, dtype=torch.float64)
    assert torch.allclose(r @ r.transpose(-1, -2), eye.expand_as(r), atol=1e-12)
    assert torch.allclose(torch.linalg.det(r), torch.ones(50, dtype=torch.float64), atol=1e-12)


def test_rotation_theta_zero_reduces_to_rz():
    # theta = 0: Rz(phi) Rz(psi) = Rz(phi + psi)
    r = og.euler_to_rotmat_zyz(torch.tensor([[30.0, 0.0, 45.0]]))[0]
    a = math.radians(75.0)
    ref = torch.tensor([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]],
                       dtype=torch.float64)
    assert torch.allclose(r, ref, atol=1e-12)


def test_rotation_known_quarter_turns():
    # theta = 90 about y maps z -> x ; phi = 90 about z maps x -> y
    ry = og.euler_to_rotmat_zyz(torch.tensor([[0.0, 90.0, 0.0]]))[0]
    assert torch.allclose(ry @ torch.tensor([0.0, 0, 1], dtype=torch.float64),
                          torch.tensor([1.0, 0, 0], dtype=torch.float64), atol=1e-12)
    rz = og.euler_to_rotmat_zyz(torch.tensor([[90.0, 0.0, 0.0]]))[0]
    assert torch.allclose(rz @ torch.tensor([1.0, 0, 0], dtype=torch.float64),
                          torch.tensor([0.0, 1, 0], dtype=torch.float64), atol=1e-12)


def test_rotmat_to_euler_roundtrip():
    eul = torch.tensor(np.random.default_rng(1).uniform([0, 1, 0], [360, 179, 360], (40, 3)))
    back = og.rotmat_to_euler_zyz(og.euler_to_rotmat_zyz(eul))
    assert torch.allclose(og.euler_to_rotmat_zyz(back), og.euler_to_rotmat_zyz(eul), atol=1e-10)
    assert torch.all(back[:, 1] >= 0) and torch.all(back[:, 1] <= 180)


def _rand_volume_rfft(n, seed=0):
    g = torch.Generator().manual_seed(seed)
    re = torch.randn((n, n, n // 2 + 1), generator=g, dtype=torch.float64)
    im = torch.randn((n, n, n // 2 + 1), generator=g, dtype=torch.float64)
    return torch.complex(re, im)


def test_slice_identity_is_centre_plane():
    n = 16
    vol = _rand_volume_rfft(n)
    s = og.extract_central_slices_rfft_3d(vol, torch.eye(3, dtype=torch.float64)[None])[0]
    assert torch.equal(s, vol[n // 2])


def test_slice_quarter_turn_about_z_is_plane_permutation():
    # R = Rz(90): (x, y) -> (-y, x); samples land on grid points (exact values)
    n = 16
    vol = _rand_volume_rfft(n, 1)
    r = og.euler_to_rotmat_zyz(torch.tensor([[90.0, 0.0, 0.0]]))
    s = og.extract_central_slices_rfft_3d(vol, r)[0]
    c = n // 2
    for iy in range(n):
        for kx in range(n // 2 + 1):
            fy = iy - c
            x, y = -fy, kx  # rotated frequency in samples
            conj = x < 0
            if conj:
                x, y = -x, -y
            yi = y + c
            inside = 0 <= yi <= n - 1 and 0 <= x <= n // 2
            exp = vol[c, yi, x] if inside else torch.zeros((), dtype=vol.dtype)
            if conj:
                exp = exp.conj()
            assert torch.allclose(s[iy, kx], exp, atol=1e-12), (iy, kx)


def test_trilinear_matches_scipy_map_coordinates():
    rng = np.random.default_rng(3)
    vol = rng.standard_normal((8, 9, 5))
    coords = rng.uniform(-0.5, [7.5, 8.5, 4.5], (200, 3))
    got = og.trilinear_sample_zero_outside(torch.tensor(vol), torch.tensor(coords)).numpy()
    ref = ndimage.map_coordinates(vol, coords.T, order=1, mode="nearest")
    inside = np.all((coords >= 0) & (coords <= np.array(vol.shape) - 1), axis=1)
    ref = np.where(inside, ref, 0.0)
    assert np.allclose(got, ref, atol=1e-12)


def test_slice_random_rotation_matches_bruteforce_float64():
    n = 8
    vol = _rand_volume_rfft(n, 2)
    rot = og.euler_to_rotmat_zyz(torch.tensor([[33.0, 71.0, 128.0]]))
    s = og.extract_central_slices_rfft_3d(vol, rot)[0].numpy()
    v = vol.numpy()
    r = rot[0].numpy()
    for iy in range(n):
        for kx in range(n // 2 + 1):
            f = np.array([kx / n, (iy - n // 2) / n, 0.0])
            g = r @ f
            conj = g[0] < 0
            if conj:
                g = -g
            c = np.array([g[2] * n + n // 2, g[1] * n + n // 2, g[0] * n])
            if np.all(c >= 0) and np.all(c <= np.array(v.shape) - 1):
                val = ndimage.map_coordinates(v.real, c[:, None], order=1, mode="nearest")[0] + \
                    1j * ndimage.map_coordinates(v.imag, c[:, None], order=1, mode="nearest")[0]
            else:
                val = 0.0
            if conj:
                val = np.conj(val)
            assert abs(s[iy, kx] - val) < 1e-12


def test_fourier_slice_theorem_gaussian_blobs():
    """irfft2 of a central slice equals the analytic projection of the blob volume."""
    from leopard_em_amd import synthetic
    from oracle import filters as of

    t = synthetic.make_blob_template(0, 48, 60, 14.0)
    vol = t.volume(dtype=torch.float64)
    vdft = of.volume_to_rfft_fourier_slice(vol)
    eul = np.array([[40.0, 60.0, 100.0]])
    rot = og.euler_to_rotmat_zyz(torch.tensor(eul))
    sl = og.extract_central_slices_rfft_3d(vdft, rot)[0]
    proj = torch.fft.ifftshift(torch.fft.irfft2(torch.fft.ifftshift(sl, dim=(0,))), dim=(0, 1))
    ana = synthetic.render_projections(t, eul, np.array([[24.0, 24.0]]), (48, 48))
    a = (proj - proj.mean()).flatten()
    b = (ana - ana.mean()).flatten()
    corr = float(a @ b / (a.norm() * b.norm()))
    assert corr > 0.99


def test_uniform_grid_count_pinned_by_reference_test():
    # tests/pydantic_models/test_orientation_search.py:99-110 -> 24 orientations
    assert og.get_uniform_euler_angles(psi_step=90.0, theta_step=90.0).shape == (24, 3)


def test_uniform_grid_ranges():
    e = og.get_uniform_euler_angles(psi_step=5.0, theta_step=10.0)
    assert e.dtype == torch.float32
    assert float(e[:, 0].min()) >= 0 and float(e[:, 0].max()) < 360
    assert float(e[:, 1].min()) == 0 and float(e[:, 1].max()) == 180
    assert float(e[:, 2].max()) < 360
    assert 25000 < e.shape[0] < 35000  # ~412 sphere points x 72 psi


def test_search_tensors_pinned_by_reference_tests():
    # tests/pydantic_models/test_defocus_search.py:25-63 and :96-108
    assert torch.allclose(og.get_search_tensors(-1000.0, 1000.0, 100.0),
                          torch.arange(-1000, 1001, 100, dtype=torch.float32))
    assert torch.allclose(og.get_search_tensors(-400.0, 400.0, 200.0),
                          torch.tensor([-400, -200, 0, 200, 400], dtype=torch.float32))
    # no zero on the grid -> 0 inserted and sorted; arange(min, max + step) may pass max
    v = og.get_search_tensors(-250.0, 250.0, 200.0)
    assert torch.allclose(v, torch.tensor([-250, -50, 0, 150, 350], dtype=torch.float32))


def test_cs_range_closed_form():
    cs = og.get_cs_range(1.0, torch.tensor([-0.1, 0.0, 0.1]), 2.7)
    assert torch.allclose(cs, torch.tensor([2.7 / 0.9**4, 2.7, 2.7 / 1.1**4]))


@pytest.mark.parametrize("seed", [0, 1])
def test_combine_euler_angles_composition(seed):
    rng = np.random.default_rng(seed)
    a = torch.tensor(rng.uniform([0, 5, 0], [360, 175, 360]), dtype=torch.float64)
    b = torch.tensor(rng.uniform([0, 5, 0], [360, 175, 360]), dtype=torch.float64)
    c = og.combine_euler_angles(a, b)
    assert torch.allclose(og.euler_to_rotmat_zyz(c), og.euler_to_rotmat_zyz(a) @ og.euler_to_rotmat_zyz(b),
                          atol=1e-10)


# ------------------------------------------------------------ anisotropic magnification
def _linear_slice(n, coef):
    """Complex slice whose value is linear in the frequency: c0 + c1 x + c2 y (in samples)."""
    fy = torch.arange(n, dtype=torch.float64) - n // 2
    fx = torch.arange(n // 2 + 1, dtype=torch.float64)
    gy, gx = torch.meshgrid(fy, fx, indexing="ij")
    return coef[0] + coef[1] * gx + coef[2] * gy


def test_transform_slice_identity_is_exact():
    n = 16
    vol = _rand_volume_rfft(n, 5)
    s = og.extract_central_slices_rfft_3d(vol, og.euler_to_rotmat_zyz(torch.tensor([[10.0, 50.0, 200.0]])))
    assert torch.equal(og.transform_slice_2d(s, torch.eye(2)), s)


def test_transform_slice_quarter_turn_is_grid_permutation():
    # A = [[0, -1], [1, 0]]: (x, y) -> (-y, x) lands on grid points; fold for x' < 0
    n = 16
    g = torch.Generator().manual_seed(6)
    s = torch.complex(torch.randn(n, n // 2 + 1, generator=g, dtype=torch.float64),
                      torch.randn(n, n // 2 + 1, generator=g, dtype=torch.float64))
    t = og.transform_slice_2d(s, torch.tensor([[0.0, -1.0], [1.0, 0.0]]))
    c = n // 2
    for iy in range(n):
        for kx in range(n // 2 + 1):
            x, y = -(iy - c), kx
            conj = x < 0
            if conj:
                x, y = -x, -y
            inside = 0 <= y + c <= n - 1 and 0 <= x <= n // 2
            exp = s[y + c, x] if inside else torch.zeros((), dtype=s.dtype)
            assert torch.allclose(t[iy, kx], exp.conj() if conj else exp, atol=1e-14), (iy, kx)


def test_transform_slice_reproduces_linear_functions():
    # bilinear interpolation is exact on a linear function, so at every output point whose
    # transformed frequency lies inside the grid the result is the function's value there
    # (conjugated where the Hermitian fold applies); outside it is zero
    n = 16
    coef = (0.3 + 0.1j, -0.7 + 0.2j, 0.25 - 0.4j)
    s = _linear_slice(n, coef)
    a = torch.tensor([[1.03, 0.08], [-0.05, 0.97]], dtype=torch.float64)
    t = og.transform_slice_2d(s, a)
    c = n // 2
    for iy in range(n):
        for kx in range(n // 2 + 1):
            k = a.numpy() @ np.array([kx, iy - c], dtype=np.float64)
            conj = k[0] < 0
            if conj:
                k = -k
            inside = 0 <= k[1] + c <= n - 1 and 0 <= k[0] <= n // 2
            exp = coef[0] + coef[1] * k[0] + coef[2] * k[1] if inside else 0.0
            if conj:
                exp = np.conj(exp)
            assert abs(complex(t[iy, kx]) - exp) < 1e-12, (iy, kx)


def test_transform_slice_matches_scipy_bilinear():
    n = 12
    rng = np.random.default_rng(8)
    s = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
    a = np.array([[0.9, 0.2], [-0.15, 1.1]])
    t = og.transform_slice_2d(torch.tensor(s), torch.tensor(a)).numpy()
    for iy in range(n):
        for kx in range(n // 2 + 1):
            k = a @ np.array([kx / n, (iy - n // 2) / n])
            conj = k[0] < 0
            if conj:
                k = -k
            cyx = np.array([k[1] * n + n // 2, k[0] * n])
            if np.all(cyx >= 0) and np.all(cyx <= np.array(s.shape) - 1):
                val = ndimage.map_coordinates(s.real, cyx[:, None], order=1, mode="nearest")[0] + \
                    1j * ndimage.map_coordinates(s.imag, cyx[:, None], order=1, mode="nearest")[0]
            else:
                val = 0.0
            assert abs(t[iy, kx] - (np.conj(val) if conj else val)) < 1e-12
