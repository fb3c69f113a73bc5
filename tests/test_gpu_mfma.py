"""Matrix-core column transforms (plan option fft_engine: 1 fp16 and 2 bf16 with the data
split into hi / lo parts, 3 fp16 and 4 bf16 with the data rounded once per stage) against
the float64 oracle.

Bounds (the precision variant's: split hi / lo DFT matrices, FP32 accumulation):
* MIP within 1e-3 of max |MIP| of the float64 oracle;
* correlation mean and std within 1e-3 relative;
* scaled MIP within 1e-3 in the conditioned form of ``test_gpu_match.z_error``;
* best orientation / defocus identical wherever the float64 top-1 / top-2 margin exceeds
  1e-3 max |MIP|;
* z-score peak list identical for a cut-off at least 1e-3 away from every float64 z.
The errors of every output are printed next to those of the FP32 path.
"""

import os
import sys

import pytest
import torch

from leopard_em_amd import core_match_template

from oracle import peaks as op
from oracle import search as osr
from test_gpu_match import _inputs, _top2_margin, z_error

pytestmark = pytest.mark.gpu

TOL = 1e-3
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, device, B, engine):
    old = os.environ.get("TDTM_FFT_ENGINE")
    os.environ["TDTM_FFT_ENGINE"] = str(engine)
    try:
        return core_match_template(*args, device=device, orientation_batch_size=B)
    finally:
        if old is None:
            del os.environ["TDTM_FFT_ENGINE"]
        else:
            os.environ["TDTM_FFT_ENGINE"] = old


def _errors(got, r64, margin, tol):
    scale = float(r64["mip"].abs().max())
    out = {"mip": float((got["mip"].double() - r64["mip"]).abs().max()) / scale}
    for k in ("correlation_mean", "correlation_variance"):
        out[k] = float((got[k].double() - r64[k]).abs().max() / r64[k].abs().max())
    out["scaled_mip"], out["scaled_mip_plain"] = z_error(got["scaled_mip"], r64["scaled_mip"],
                                                         r64["correlation_variance"], scale)
    ok = margin > tol * scale
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        out[k + "_px"] = int((((got[k] != r64[k].float()) & ok)).sum())
    return out


def _check(errs, tol):
    for k in ("mip", "correlation_mean", "correlation_variance", "scaled_mip"):
        assert errs[k] < tol, (k, errs[k])
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        assert errs[k + "_px"] == 0, k


def _peaks_identical(got, r64, cut, valid):
    z = r64["scaled_mip"][:valid[0], :valid[1]]
    while bool(((z - cut).abs() < TOL * float(z.abs().max())).any()):
        cut += 0.01
    gy, gx = op.find_peaks_from_zscore(got["scaled_mip"][:valid[0], :valid[1]], cut)
    ry, rx = op.find_peaks_from_zscore(z.float(), cut)
    return cut, len(ry), gy.tolist() == ry.tolist() and gx.tolist() == rx.tolist()


@pytest.mark.parametrize("engine", [1, 2, 3, 4], ids=["fp16-split", "bf16-split", "fp16", "bf16"])
def test_mfma_column_pass_small(engine, cuda_device):
    """4096 x 1024 grid, 128^3 template, 2 defocus, 8 orientations."""
    print(f"fft_engine {engine}: 4096x1024 h128 D2 O8", flush=True)
    args = _inputs(4096, 1024, 128, 2, 1, 8, seed=77)
    got = _run(args, cuda_device, 4, engine)
    f32 = _run(args, cuda_device, 4, 0)
    r64 = osr.core_match_template(*args, orientation_batch_size=4, dtype=torch.float64)
    margin = _top2_margin(args)
    e = _errors(got, r64, margin, TOL)
    e32 = _errors(f32, r64, margin, TOL)
    print("engine errors vs fp64: " + ", ".join(f"{k} {v:.2e}" if isinstance(v, float) else f"{k} {v}"
                                              for k, v in e.items()), flush=True)
    print("fp32 path errors vs fp64: " + ", ".join(f"{k} {v:.2e}" if isinstance(v, float) else f"{k} {v}"
                                                 for k, v in e32.items()), flush=True)
    _check(e, TOL)


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("engine", [1, 2], ids=["fp16-split", "bf16-split"])
def test_mfma_column_pass_4096_t512(engine, cuda_device):
    """4096 x 4096 micrograph, 512^3 template, 13 defocus, 4 orientations (two true poses)."""
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import bench
    from leopard_em_amd.utils.search import get_uniform_euler_angles

    from test_gpu_search_4096_t512 import _margin64

    print(f"fft_engine {engine}: 4096^2 x 512^3, 13 defocus", flush=True)
    img_dft, tdft, ctf, wf, dfs, vol, image, mg = bench.make_inputs(cuda_device, 4096, 4096, 512, 5000, 200.0, 200,
                                                                    return_micrograph=True)
    del vol, image
    grid = get_uniform_euler_angles(psi_step=1.5, theta_step=2.5)
    eul = torch.cat([torch.tensor(mg.eulers[:2], dtype=torch.float32), grid[[1000, 700000]].float()])
    pix = torch.tensor([0.0])
    args_dev = (img_dft, tdft, ctf, wf, dfs, pix, eul)
    got = _run(args_dev, cuda_device, 2, engine)
    args = (img_dft.cpu(), tdft.cpu(), ctf.cpu(), wf.cpu(), dfs, pix, eul)
    del img_dft, tdft, ctf, wf, args_dev
    torch.cuda.empty_cache()
    r64 = osr.core_match_template(*args, orientation_batch_size=2, dtype=torch.float64)
    filt = args[2].double() * args[3].double()[None, None]
    margin = _margin64(args[0], args[1], filt, eul, 2)
    e = _errors(got, r64, margin, TOL)
    print("engine errors vs fp64: " + ", ".join(f"{k} {v:.2e}" if isinstance(v, float) else f"{k} {v}"
                                              for k, v in e.items()), flush=True)
    _check(e, TOL)
    # 52 correlations per pixel bound z by 51 / sqrt(52) ~ 7.1: the cut-off sits below the
    # oracle's largest z
    zmax = float(r64["scaled_mip"][:4096 - 511, :4096 - 511].max())
    cut, n, same = _peaks_identical(got, r64, 0.9 * zmax, (4096 - 511, 4096 - 511))
    print(f"peak list at z > {cut:.2f}: {n} peaks, identical: {same}", flush=True)
    assert n >= 1 and same
