"""GPU parity at the production search shape: 4096 x 4096 micrograph, 512^3 template,
13 defocus planes (-1200..+1200 A step 200), orientation batches of 2.

Inputs follow the benchmark's synthetic recipe (``bench.make_inputs``: 5000-blob
template, 200 embedded CTF-filtered projections with per-particle defocus, SNR 0.05,
whitening from the image's radial power spectrum).  Four orientations -- two true poses
of embedded particles and two from the 2.5 / 1.5 degree grid -- give 52 correlations,
which the float32 and float64 CPU oracles run in seconds.

Bounds (the same as the smaller parity cases, ``test_gpu_match``):
* MIP within 1e-4 of max |MIP| against the float32 and the float64 oracle;
* correlation mean and std within 1e-4 relative of the float64 oracle;
* scaled MIP (z): error propagated from the MIP bound, |dz| std / (max |MIP| (1 + |z|)),
  below 1e-4 against the float64 oracle (``test_gpu_match.z_error``);
* best phi / theta / psi / defocus identical wherever the float64 top-1 / top-2 margin
  exceeds 1e-4 max |MIP|.
The error of every output against the float64 oracle is printed.
"""

import os
import sys

import pytest
import torch

from leopard_em_amd import core_match_template

from oracle import geometry as og
from oracle import search as osr
from test_gpu_match import z_error

pytestmark = pytest.mark.gpu

TOL = 1e-4
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _margin64(img_dft, tdft, filt, eul, batch):
    """Float64 top-1 minus top-2 correlation over all (defocus, orientation) per pixel."""
    X = img_dft.to(torch.complex128)
    T = tdft.to(torch.complex128)
    top1 = top2 = None
    for i in range(0, eul.shape[0], batch):
        rot = og.euler_to_rotmat_zyz(eul[i:i + batch], torch.float64)
        cc = osr.batched_cross_correlate(X, T, rot, filt)
        cc = cc.reshape(-1, *cc.shape[-2:])
        if top1 is not None:
            cc = torch.cat([top1[None], top2[None], cc])
        t = torch.topk(cc, 2, dim=0).values
        top1, top2 = t[0].clone(), t[1].clone()
        del cc, t
    return top1 - top2


@pytest.mark.timeout(1200)
def test_search_parity_4096_t512_d13(cuda_device):
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import bench
    from leopard_em_amd.utils.search import get_uniform_euler_angles

    print("4096^2 x 512^3 parity: building inputs", flush=True)
    img_dft, tdft, ctf, wf, dfs, vol, image, mg = bench.make_inputs(cuda_device, 4096, 4096, 512, 5000, 200.0, 200,
                                                                    return_micrograph=True)
    del vol, image
    grid = get_uniform_euler_angles(psi_step=1.5, theta_step=2.5)
    eul = torch.cat([torch.tensor(mg.eulers[:2], dtype=torch.float32), grid[[1000, 700000]].float()])
    pix = torch.tensor([0.0])
    B = 2
    got = core_match_template(img_dft, tdft, ctf, wf, dfs, pix, eul, device=cuda_device, orientation_batch_size=B)
    print("GPU search done; float32 oracle ...", flush=True)
    args = (img_dft.cpu(), tdft.cpu(), ctf.cpu(), wf.cpu(), dfs, pix, eul)
    del img_dft, tdft, ctf, wf
    torch.cuda.empty_cache()
    r32 = osr.core_match_template(*args, orientation_batch_size=B)
    e32 = float((got["mip"].double() - r32["mip"].double()).abs().max())
    best32 = {k: r32[k] for k in ("best_phi", "best_theta", "best_psi", "best_defocus")}
    del r32
    print("float64 oracle ...", flush=True)
    r64 = osr.core_match_template(*args, orientation_batch_size=B, dtype=torch.float64)
    scale = float(r64["mip"].abs().max())
    e64 = float((got["mip"].double() - r64["mip"]).abs().max()) / scale
    e32 /= scale
    print(f"mip: rel err vs fp32 {e32:.2e}, vs fp64 {e64:.2e} (max |mip| {scale:.3f})", flush=True)
    assert e32 < TOL and e64 < TOL
    errs = {"mip": e64}
    for k in ("correlation_mean", "correlation_variance"):
        ref = r64[k]
        err = float((got[k].double() - ref).abs().max() / ref.abs().max())
        errs[k] = err
        assert err < TOL, k
    ez, ezr = z_error(got["scaled_mip"], r64["scaled_mip"], r64["correlation_variance"], scale)
    errs["scaled_mip (conditioned)"] = ez
    errs["scaled_mip (plain relative)"] = ezr
    assert ez < TOL
    print("float64 top-2 margins ...", flush=True)
    filt = args[2].double() * args[3].double()[None, None]
    margin = _margin64(args[0], args[1], filt, eul, B)
    ok = margin > TOL * scale
    print(f"{int(ok.sum())} of {ok.numel()} pixels with a margin above {TOL} max|mip|", flush=True)
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        agree = (got[k] == best32[k]) | ~ok
        agree64 = (got[k] == r64[k].float()) | ~ok
        errs[k] = int((~agree64).sum())
        assert bool(agree.all()) and bool(agree64.all()), f"{k}: {int((~agree).sum())} pixels differ"
    print("errors vs fp64 oracle: " + ", ".join(f"{k} {v:.2e}" if isinstance(v, float) else f"{k} {v} px"
                                                for k, v in errs.items()), flush=True)
    assert got["total_projections"] == 4 * 13
