"""GPU tests of the search harness: checkpoint / resume and the YAML + MRC manager run."""

import os

import numpy as np
import pytest
import torch
import yaml

from leopard_em_amd import synthetic
from leopard_em_amd.backend.core_match_template import core_match_template, core_match_template_resumable
from leopard_em_amd.io import mrc
from leopard_em_amd.manager import MatchTemplateManager
from oracle import filters as of
from oracle import peaks as op
from oracle import search as osr

pytestmark = pytest.mark.gpu


def _search_inputs(H, W, h, D, O, seed):
    t = synthetic.make_blob_template(seed, h, 150, 0.35 * h)
    mg = synthetic.make_micrograph(t, seed, (H, W), 3, 12200.0, 11800.0, 30.0, snr=0.5)
    img_dft, wf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    dfs = torch.linspace(-200.0, 200.0, D)
    pix = torch.tensor([0.0])
    ctf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0).float()
    eul = torch.tensor(synthetic.random_uniform_rotations(np.random.default_rng(seed), O), dtype=torch.float32)
    return (img_dft.to(torch.complex64), of.volume_to_rfft_fourier_slice(t.volume()).to(torch.complex64), ctf,
            wf.float(), dfs, pix, eul)


def test_resume_is_bit_exact(tmp_path, cuda_device):
    args = _search_inputs(256, 256, 64, 3, 22, seed=11)
    print("resume: 256^2, h 64, D 3, O 22, B 4", flush=True)
    full = core_match_template(*args, device=cuda_device, orientation_batch_size=4)
    ck = str(tmp_path / "stats.npz")
    assert core_match_template_resumable(*args, device=cuda_device, checkpoint_path=ck, orientation_batch_size=4,
                                         max_batches=2) is None
    assert core_match_template_resumable(*args, device=cuda_device, checkpoint_path=ck, orientation_batch_size=4,
                                         max_batches=3) is None
    res = core_match_template_resumable(*args, device=cuda_device, checkpoint_path=ck, orientation_batch_size=4)
    for k, v in full.items():
        if isinstance(v, torch.Tensor):
            assert torch.equal(v, res[k]), k
        else:
            assert v == res[k], k
    # a checkpoint of another search is refused
    other = list(args)
    other[6] = args[6].flip(0).contiguous()
    with pytest.raises(ValueError, match="different search"):
        core_match_template_resumable(*other, device=cuda_device, checkpoint_path=ck, orientation_batch_size=4)


@pytest.mark.parametrize("H,W", [(256, 256), (250, 300)])
def test_manager_yaml_mrc_end_to_end(tmp_path, cuda_device, H, W):
    """YAML + MRC in, exported maps and peak table out; (250, 300) runs the device
    preparation's direct DFTs and the search in a zero-padded 256 x 512 grid."""
    h = 32
    t = synthetic.make_blob_template(21, h, 120, 11.0)
    mg = synthetic.make_micrograph(t, 21, (H, W), 3, 12200.0, 11800.0, 30.0, snr=0.5)
    mrc.write_mrc_from_tensor(mg.image.float(), tmp_path / "micrograph.mrc", voxel_size=1.0)
    mrc.write_mrc_from_tensor(t.volume().float(), tmp_path / "template.mrc", voxel_size=1.0)
    names = ["mip_path", "scaled_mip_path", "orientation_psi_path", "orientation_theta_path", "orientation_phi_path",
             "relative_defocus_path", "correlation_average_path", "correlation_variance_path"]
    cfg = {
        "template_volume_path": str(tmp_path / "template.mrc"),
        "micrograph_path": str(tmp_path / "micrograph.mrc"),
        "match_template_result": dict({k: str(tmp_path / f"{k}.mrc") for k in names}, allow_file_overwrite=True),
        "optics_group": {"label": "s", "voltage": 300.0, "pixel_size": 1.0, "defocus_u": 12200.0,
                         "defocus_v": 11800.0, "astigmatism_angle": 30.0},
        "defocus_search_config": {"defocus_min": -200.0, "defocus_max": 200.0, "defocus_step": 200.0},
        "orientation_search_config": {"psi_step": 60.0, "theta_step": 60.0, "symmetry": "C1"},
        "preprocessing_filters": {"whitening_filter": {"enabled": True}},
        "computational_config": {"gpu_ids": [0], "num_cpus": 1},
    }
    p = tmp_path / "cfg.yaml"
    p.write_text(yaml.safe_dump(cfg))
    m = MatchTemplateManager.from_yaml(p)
    kw = m.make_backend_core_function_kwargs()
    O = kw["euler_angles"].shape[0]
    print(f"manager: {H}x{W}, h {h}, D 3, O {O}", flush=True)
    assert m.run_match_template(orientation_batch_size=8) is True
    r = m.match_template_result
    # the manager builds its inputs on the GPU; the oracle search runs on CPU copies
    assert kw["image_dft"].is_cuda and kw["template_dft"].is_cuda and kw["ctf_filters"].is_cuda
    ck = {k: (v.cpu() if isinstance(v, torch.Tensor) else v) for k, v in kw.items()}
    ref = osr.core_match_template(ck["image_dft"], ck["template_dft"], ck["ctf_filters"],
                                  ck["whitening_filter_template"], ck["defocus_values"], ck["pixel_values"],
                                  ck["euler_angles"], orientation_batch_size=O)
    crop = (slice(0, H - h + 1), slice(0, W - h + 1))
    scale = float(ref["mip"].abs().max())
    assert float((r.mip - ref["mip"][crop]).abs().max()) / scale < 1e-4
    assert r.mip.shape == (H - h + 1, W - h + 1)
    for k, pth in zip(["mip", "scaled_mip", "orientation_psi", "orientation_theta", "orientation_phi",
                       "relative_defocus", "correlation_average", "correlation_variance"], names):
        assert torch.equal(mrc.load_mrc_image(getattr(r, pth)), getattr(r, k)), k
    # a coarse 60-degree grid misses the particles' poses, so pick with an explicit cutoff
    cut = 0.8 * float(r.scaled_mip.max())
    df = m.results_to_dataframe(locate_peaks_kwargs={"z_score_cutoff": cut})
    oy, ox = op.find_peaks_from_zscore(r.scaled_mip, cut)
    assert len(df) == len(oy) >= 1
    assert df["pos_y"].tolist() == oy.tolist() and df["pos_x"].tolist() == ox.tolist()
    assert (df["scaled_mip"] > cut).all()
    # checkpointed run through the manager gives the same maps
    m2 = MatchTemplateManager.from_yaml(p)
    assert m2.run_match_template(orientation_batch_size=8, do_result_export=False,
                                 checkpoint_path=str(tmp_path / "ck.npz")) is True
    assert torch.equal(m2.match_template_result.mip, r.mip)
    assert torch.equal(m2.match_template_result.orientation_phi, r.orientation_phi)
    assert os.path.exists(tmp_path / "ck.npz")


def test_autotune_keeps_results_and_caches(tmp_path, cuda_device):
    from leopard_em_amd.backend import MatchPlan, raw_statistics
    from leopard_em_amd.backend.autotune import autotune

    args = _search_inputs(512, 512, 64, 2, 16, seed=13)
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    tdft, eul = tdft.to(cuda_device), eul.to(cuda_device)
    ref = MatchPlan(512, 512, 64, 4, 2, 1, cuda_device)
    ref.set_inputs(img_dft, ctf, wf)
    ref.reset_stats()
    ref.run(tdft, eul, 0, 16)
    plan = MatchPlan(512, 512, 64, 4, 2, 1, cuda_device)
    plan.set_inputs(img_dft, ctf, wf)
    plan.reset_stats()
    plan.run(tdft, eul, 0, 8)
    cache = str(tmp_path / "tune.json")
    print("autotune 512^2 h64 B4 D2", flush=True)
    res = autotune(plan, tdft, eul, steps=2, cache_path=cache)
    assert not res["cached"] and res["ms_per_step"] > 0
    assert any("ms_per_step" in t for t in res["timings"])
    plan.run(tdft, eul, 8, 16)  # statistics were restored: the search continues unchanged
    torch.cuda.synchronize()
    a, b = raw_statistics(ref), raw_statistics(plan)
    for k in a:
        assert torch.equal(a[k], b[k]), k
    again = autotune(plan, tdft, eul, cache_path=cache)
    assert again["cached"] and again["options"] == res["options"]
