"""Seeded synthetic inputs for 2DTM: Gaussian-blob templates and micrographs.

The template volume is a sum of isotropic 3D Gaussians, so every projection of it is
an analytic sum of 2D Gaussians; micrographs are rendered from those analytic
projections (no Fourier-slice code involved), filtered by a CTF and given Gaussian
noise at a stated SNR = var(signal) / var(noise).

Geometry convention (matches the Fourier central-slice convention of
``torch_fourier_slice`` used by the reference, ``backend/cross_correlation.py:76``):
a slice taken with rotation R is the projection along z of v(R s), so a blob at
offset c (xyz, voxels from the volume centre) appears at (R^T c)_xy pixels from the
projection centre.  A particle whose projection centre sits at image pixel (Y, X)
correlates best at the top-left position (Y - h//2, X - w//2).

Seeds: volume uses ``seed``, poses/positions ``seed + 1``, noise ``seed + 2``
(numpy PCG64).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch


def _rotmat_zyz_np(phi, theta, psi):
    a, b, c = np.radians(phi), np.radians(theta), np.radians(psi)
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rz1 = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz2 = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rz1 @ ry @ rz2


def _rotmat_to_zyz_np(r):
    theta = math.degrees(math.acos(max(-1.0, min(1.0, r[2, 2]))))
    if abs(math.sin(math.radians(theta))) < 1e-9:
        phi = math.degrees(math.atan2(r[1, 0], r[0, 0]))
        psi = 0.0
    else:
        phi = math.degrees(math.atan2(r[1, 2], r[0, 2]))
        psi = math.degrees(math.atan2(r[2, 1], -r[2, 0]))
    return phi % 360.0, theta, psi % 360.0


def random_uniform_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform SO(3) samples as ZYZ Euler angles (degrees), from normalised N(0, I4) quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    out = np.empty((n, 3))
    for i, (w, x, y, z) in enumerate(q):
        r = np.array([
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ])
        out[i] = _rotmat_to_zyz_np(r)
    return out


@dataclass
class BlobTemplate:
    """A Gaussian-blob template: centres (xyz, voxels from the volume centre) and amplitudes."""

    size: int
    centers: np.ndarray
    amplitudes: np.ndarray
    sigma: float = 1.5

    def volume(self, device="cpu", dtype=torch.float32) -> torch.Tensor:
        """Render the (size, size, size) volume, index order [z, y, x], centre at size//2."""
        n = self.size
        r = int(math.ceil(5 * self.sigma))
        vol = torch.zeros((n + 2 * r, n + 2 * r, n + 2 * r), dtype=torch.float64, device=device)
        off = torch.arange(-r, r + 1, dtype=torch.float64, device=device)
        cen = torch.as_tensor(self.centers, dtype=torch.float64, device=device) + n // 2
        amp = torch.as_tensor(self.amplitudes, dtype=torch.float64, device=device)
        base = torch.round(cen).long()
        frac = cen - base.to(torch.float64)
        g = torch.exp(-((off[None, None, :] - frac[:, :, None]) ** 2) / (2 * self.sigma**2))  # (B, 3, 2r+1)
        for i in range(cen.shape[0]):
            bx, by, bz = (int(v) for v in base[i])
            blk = amp[i] * g[i, 2][:, None, None] * g[i, 1][None, :, None] * g[i, 0][None, None, :]
            vol[bz:bz + 2 * r + 1, by:by + 2 * r + 1, bx:bx + 2 * r + 1] += blk
        return vol[r:r + n, r:r + n, r:r + n].to(dtype).contiguous()


def make_blob_template(seed: int, size: int, n_blobs: int, radius: float, sigma: float = 1.5) -> BlobTemplate:
    """Blob centres uniform in a ball of ``radius`` voxels, amplitudes U(0.5, 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.standard_normal((n_blobs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = radius * rng.random(n_blobs) ** (1.0 / 3.0)
    centers = d * rad[:, None]
    amps = rng.uniform(0.5, 1.0, n_blobs)
    return BlobTemplate(size, centers, amps, sigma)


def render_projections(template: BlobTemplate, eulers: np.ndarray, positions: np.ndarray,
                       shape: tuple[int, int], device="cpu") -> torch.Tensor:
    """Sum of analytic projections: particle i centred at pixel positions[i] = (Y, X)."""
    H, W = shape
    s = template.sigma
    r = int(math.ceil(5 * s))
    img = torch.zeros(H * W, dtype=torch.float64, device=device)
    off = torch.arange(-r, r + 1, device=device)
    amp2d = torch.as_tensor(template.amplitudes * math.sqrt(2 * math.pi) * s, device=device)
    cen = torch.as_tensor(template.centers, dtype=torch.float64, device=device)
    for (phi, theta, psi), (Y, X) in zip(eulers, positions):
        rt = torch.as_tensor(_rotmat_zyz_np(phi, theta, psi).T, device=device)
        p = cen @ rt.T  # rows: R^T c
        cy = p[:, 1] + Y
        cx = p[:, 0] + X
        by = torch.round(cy).long()
        bx = torch.round(cx).long()
        gy = torch.exp(-((by[:, None] + off[None] - cy[:, None]) ** 2) / (2 * s * s))
        gx = torch.exp(-((bx[:, None] + off[None] - cx[:, None]) ** 2) / (2 * s * s))
        vals = amp2d[:, None, None] * gy[:, :, None] * gx[:, None, :]
        iy = (by[:, None] + off[None]) % H
        ix = (bx[:, None] + off[None]) % W
        flat = (iy[:, :, None] * W + ix[:, None, :]).reshape(-1)
        img.index_add_(0, flat, vals.reshape(-1))
    return img.reshape(H, W)


@dataclass
class SyntheticMicrograph:
    image: torch.Tensor
    eulers: np.ndarray
    centers_yx: np.ndarray
    defocus: np.ndarray
    meta: dict = field(default_factory=dict)


def ctf_2d_np(shape, pixel_size, defocus_a, astig_a, astig_angle_deg, voltage_kv=300.0,
              cs_mm=2.7, amplitude_contrast=0.07, phase_shift_deg=0.0):
    """CTF on the unshifted rfft grid of a ``shape`` image (torch_ctf sign convention).

    CTF(k) = -sin(-pi lam df(a) k^2 + pi/2 Cs lam^3 k^4 - atan(A/sqrt(1-A^2)) - phi),
    df(a) = df + astig cos(2 (a - a_ast)), a = atan2(ky, kx).
    """
    h, w = shape
    v = voltage_kv * 1e3
    lam = 6.62607015e-34 / math.sqrt(2 * 9.1093837015e-31 * 1.602176634e-19 * v *
                                     (1 + 1.602176634e-19 * v / (2 * 9.1093837015e-31 * 299792458.0**2))) * 1e10
    ky = np.fft.fftfreq(h)[:, None] / pixel_size
    kx = np.fft.rfftfreq(w)[None, :] / pixel_size
    n4 = ky**2 + kx**2
    a = np.arctan2(ky, kx)
    df = defocus_a + astig_a * np.cos(2 * (a - math.radians(astig_angle_deg)))
    w_ac = math.atan(amplitude_contrast / math.sqrt(1 - amplitude_contrast**2))
    chi = -math.pi * lam * n4 * df + math.pi / 2 * cs_mm * 1e7 * lam**3 * n4**2 - w_ac - math.radians(phase_shift_deg)
    return -np.sin(chi)


def make_micrograph(template: BlobTemplate, seed: int, shape: tuple[int, int], n_particles: int,
                    defocus_u: float, defocus_v: float, astig_angle: float, snr: float,
                    pixel_size: float = 1.0, defocus_spread: float = 0.0, margin: int | None = None,
                    device="cpu") -> SyntheticMicrograph:
    """Micrograph of ``n_particles`` analytic projections, CTF-filtered, plus noise.

    Particle signal enters with negative contrast, -(projection * CTF), so that the
    contrast-flipped, CTF-filtered template projection correlates positively.
    ``defocus_spread`` > 0 gives each particle its own mean defocus,
    (u+v)/2 + U(-spread, +spread), rendered in a local box.
    """
    H, W = shape
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    eul = random_uniform_rotations(rng, n_particles)
    if margin is None:
        margin = template.size // 2
    ys = rng.uniform(margin, H - margin, n_particles)
    xs = rng.uniform(margin, W - margin, n_particles)
    pos = np.stack([ys, xs], 1)
    mean_df = 0.5 * (defocus_u + defocus_v)
    astig = abs(defocus_u - defocus_v) / 2
    dfs = mean_df + (rng.uniform(-defocus_spread, defocus_spread, n_particles) if defocus_spread > 0
                     else np.zeros(n_particles))
    signal = torch.zeros((H, W), dtype=torch.float64, device=device)
    if defocus_spread <= 0:
        proj = render_projections(template, eul, pos, shape, device)
        ctf = torch.as_tensor(ctf_2d_np(shape, pixel_size, mean_df, astig, astig_angle), device=device)
        signal = torch.fft.irfft2(torch.fft.rfft2(proj) * -ctf, s=shape)
    else:
        box = 1 << int(math.ceil(math.log2(template.size + 400)))
        ctf_cache = {}
        for i in range(n_particles):
            cy, cx = int(round(pos[i, 0])), int(round(pos[i, 1]))
            local = render_projections(template, eul[i:i + 1],
                                       np.array([[pos[i, 0] - cy + box // 2, pos[i, 1] - cx + box // 2]]),
                                       (box, box), device)
            key = round(float(dfs[i]), 3)
            if key not in ctf_cache:
                ctf_cache[key] = torch.as_tensor(ctf_2d_np((box, box), pixel_size, dfs[i], astig, astig_angle),
                                                 device=device)
            loc = torch.fft.irfft2(torch.fft.rfft2(local) * -ctf_cache[key], s=(box, box))
            yy = (torch.arange(box, device=device) + cy - box // 2) % H
            xx = (torch.arange(box, device=device) + cx - box // 2) % W
            signal[yy[:, None], xx[None, :]] += loc
    rng_n = np.random.Generator(np.random.PCG64(seed + 2))
    var_s = float(signal.var())
    noise = torch.as_tensor(rng_n.standard_normal((H, W)) * math.sqrt(var_s / snr), device=device)
    img = (signal + noise).to(torch.float32)
    return SyntheticMicrograph(img, eul, pos, dfs, dict(snr=snr, defocus_u=defocus_u, defocus_v=defocus_v,
                                                        astig_angle=astig_angle, pixel_size=pixel_size))
