"""GPU parity of the HIP match-template search against the CPU oracle.

Tolerances (float32 path):
* MIP within 1e-4 of the float32 oracle, relative to max |MIP|; error against the
  float64 oracle reported and bounded at the same level.
* correlation mean / std (reference "variance" = std) within 1e-4 relative of the
  float64 oracle (the HIP sums are compensated, the float32 oracle's are not).
* scaled MIP z = (MIP - mean) / std: its error propagated from the MIP bound,
  |dz| std / (max |MIP| (1 + |z|)), below 1e-4 against the float64 oracle (``z_error``).
  A plain relative error max |dz| / max |z| is not a usable bound: with few correlations
  per pixel some pixels have a std far below the typical one, and any rounding of the
  correlations is amplified there (the float32 oracle itself then differs from the
  float64 one by more than 1e-3).  Both figures are printed.
* best orientation / defocus identical at every pixel whose top-1 / top-2 margin
  exceeds 1e-4 * max |MIP|.
* z-score peak lists (zscore cutoff + greedy distance filter) identical.
"""

import numpy as np
import pytest
import torch

from leopard_em_amd import core_match_template, synthetic
from oracle import filters as of
from oracle import geometry as og
from oracle import peaks as op
from oracle import search as osr

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _inputs(H, W, h, D, C, O, seed, n_particles=3):
    t = synthetic.make_blob_template(seed, h, 200, 0.35 * h)
    vol = t.volume()
    mg = synthetic.make_micrograph(t, seed, (H, W), n_particles, 12200.0, 11800.0, 30.0, snr=0.5)
    img_dft, wf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    tdft = of.volume_to_rfft_fourier_slice(vol)
    dfs = torch.linspace(-200.0, 200.0, D) if D > 1 else torch.tensor([0.0])
    pix = torch.linspace(-0.02, 0.02, C) if C > 1 else torch.tensor([0.0])
    ctf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0).float()
    eul = torch.tensor(np.concatenate([mg.eulers, synthetic.random_uniform_rotations(
        np.random.default_rng(seed + 10), O - n_particles)]), dtype=torch.float32)
    return (img_dft.to(torch.complex64), tdft.to(torch.complex64), ctf, wf.float(), dfs, pix, eul)


def z_error(z, z64, std64, mip_scale):
    """(conditioned error, plain relative error) of a scaled-MIP map against the float64
    oracle.  An error e in the MIP, mean and std of a pixel moves z = (mip - mean) / std by
    about e (1 + |z|) / std, so |dz| std / (mip_scale (1 + |z|)) is the error in units of
    the MIP's own bound."""
    dz = (z.double() - z64.double()).abs()
    cond = float((dz * std64.double() / (mip_scale * (1.0 + z64.double().abs()))).max())
    return cond, float(dz.max() / z64.abs().max())


def _top2_margin(args):
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    filt = ctf.double() * wf.double()[None, None]
    rot = og.euler_to_rotmat_zyz(eul, torch.float64)
    cc = osr.batched_cross_correlate(img_dft.to(torch.complex128), tdft.to(torch.complex128), rot, filt)
    H, W = cc.shape[-2:]
    top = torch.topk(cc.reshape(-1, H, W), 2, dim=0).values
    return top[0] - top[1]


CASES = [
    # H, W, h, D, C, O, B   (column FFT = H points, row c2r = W/2-point complex FFT;
    #                        sizes >= 512 take the four-step path, smaller ones Stockham)
    (256, 256, 64, 3, 1, 16, 4),
    (256, 512, 32, 2, 2, 12, 3),
    (512, 512, 64, 1, 1, 24, 8),
    (1024, 512, 128, 2, 1, 8, 4),
    (1024, 2048, 64, 2, 1, 4, 4),
    (2048, 4096, 64, 2, 1, 4, 4),
    (4096, 1024, 128, 1, 1, 4, 2),
    # edge cases: template as large as the image, one orientation per batch with three
    # pixel sizes, a 512^3 template, many defocus planes
    (64, 64, 64, 2, 1, 4, 2),
    (128, 64, 32, 1, 3, 3, 1),
    (512, 512, 512, 1, 1, 4, 2),
    (256, 256, 32, 24, 1, 4, 4),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "H{}W{}h{}D{}C{}O{}B{}".format(*c))
def test_match_parity_against_oracle(case, cuda_device):
    H, W, h, D, C, O, B = case
    print(f"case {case}", flush=True)
    args = _inputs(H, W, h, D, C, O, seed=H + W + h)
    got = core_match_template(*args, device=cuda_device, orientation_batch_size=B)
    r32 = osr.core_match_template(*args, orientation_batch_size=B)
    r64 = osr.core_match_template(*args, orientation_batch_size=B, dtype=torch.float64)
    scale = float(r64["mip"].abs().max())
    e32 = float((got["mip"].double() - r32["mip"].double()).abs().max()) / scale
    e64 = float((got["mip"].double() - r64["mip"]).abs().max()) / scale
    print(f"mip rel err vs fp32 {e32:.2e} vs fp64 {e64:.2e}", flush=True)
    assert e32 < TOL and e64 < TOL
    for k in ("correlation_mean", "correlation_variance"):
        ref = r64[k]
        err = float((got[k].double() - ref).abs().max() / ref.abs().max())
        print(f"{k} rel err vs fp64 {err:.2e}", flush=True)
        assert err < TOL
    ez, ezr = z_error(got["scaled_mip"], r64["scaled_mip"], r64["correlation_variance"], scale)
    ez32, ezr32 = z_error(r32["scaled_mip"], r64["scaled_mip"], r64["correlation_variance"], scale)
    print(f"scaled_mip err vs fp64: conditioned {ez:.2e}, plain relative {ezr:.2e} "
          f"(fp32 oracle {ez32:.2e}, {ezr32:.2e})", flush=True)
    assert ez < TOL
    margin = _top2_margin(args)
    ok = margin > TOL * scale
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        agree = (got[k] == r32[k]) | ~ok
        assert bool(agree.all()), f"{k}: {int((~agree).sum())} pixels differ"
    assert got["total_projections"] == O * D * C


NON_POW2 = [
    # H, W, h, D, C, O, B -> search grid (next powers of two)
    (200, 300, 32, 2, 1, 8, 4),     # 256 x 512
    (100, 128, 64, 1, 1, 6, 3),     # 128 x 128 (one edge a power of two)
    (640, 960, 64, 2, 1, 4, 4),     # 1024 x 1024 (four-step transforms)
    (1000, 5000, 64, 1, 1, 4, 4),   # 1024 x 8192 (8192-sample rows)
]


@pytest.mark.parametrize("case", NON_POW2, ids=lambda c: "H{}W{}h{}D{}C{}O{}B{}".format(*c))
def test_match_non_power_of_two_image(case, cuda_device):
    """Image sizes that are not powers of two run zero-padded in the next power-of-two
    grid: the valid region (template footprint inside the image) matches the oracle's
    circular search at the image's own size."""
    H, W, h, D, C, O, B = case
    print(f"non-power-of-two case {case}", flush=True)
    args = _inputs(H, W, h, D, C, O, seed=H + 3 * W + h)
    got = core_match_template(*args, device=cuda_device, orientation_batch_size=B)
    r32 = osr.core_match_template(*args, orientation_batch_size=B)
    r64 = osr.core_match_template(*args, orientation_batch_size=B, dtype=torch.float64)
    assert tuple(got["mip"].shape) == (H, W)
    vy, vx = H - h + 1, W - h + 1
    scale = float(r64["mip"][:vy, :vx].abs().max())
    e32 = float((got["mip"][:vy, :vx].double() - r32["mip"][:vy, :vx].double()).abs().max()) / scale
    e64 = float((got["mip"][:vy, :vx].double() - r64["mip"][:vy, :vx]).abs().max()) / scale
    print(f"valid-region mip rel err vs fp32 {e32:.2e} vs fp64 {e64:.2e}", flush=True)
    assert e32 < TOL and e64 < TOL
    for k in ("correlation_mean", "correlation_variance"):
        ref = r64[k][:vy, :vx]
        err = float((got[k][:vy, :vx].double() - ref).abs().max() / ref.abs().max())
        assert err < TOL, k
    ok = (_top2_margin(args) > TOL * scale)[:vy, :vx]
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        agree = (got[k][:vy, :vx] == r32[k][:vy, :vx]) | ~ok
        assert bool(agree.all()), f"{k}: {int((~agree).sum())} pixels differ"


def test_match_partial_last_batch_exact_index(cuda_device):
    """O % B != 0 with D > 1: the HIP path decodes every correlation to its true index."""
    args = _inputs(256, 256, 32, 3, 1, 10, seed=7)
    got = core_match_template(*args, device=cuda_device, orientation_batch_size=4)
    ref = osr.core_match_template(*args, orientation_batch_size=10)  # one full batch: no quirk
    scale = float(ref["mip"].abs().max())
    assert float((got["mip"] - ref["mip"]).abs().max()) / scale < TOL
    margin = _top2_margin(args)
    ok = margin > TOL * scale
    assert bool(((got["best_defocus"] == ref["best_defocus"]) | ~ok).all())
    assert bool(((got["best_phi"] == ref["best_phi"]) | ~ok).all())


def test_match_recovers_embedded_particles(cuda_device):
    """The true orientations are in the grid: each particle's peak sits at its top-left."""
    H = W = 512
    h = 64
    t = synthetic.make_blob_template(3, h, 300, 22.0)
    vol = t.volume()
    mg = synthetic.make_micrograph(t, 3, (H, W), 4, 12200.0, 11800.0, 30.0, snr=0.2)
    img_dft, wf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    tdft = of.volume_to_rfft_fourier_slice(vol)
    dfs = torch.tensor([0.0])
    ctf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, torch.tensor([0.0]), astigmatism_angle=30.0, voltage=300.0,
        spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.0).float()
    eul = torch.tensor(np.concatenate([mg.eulers, synthetic.random_uniform_rotations(
        np.random.default_rng(99), 60)]), dtype=torch.float32)
    got = core_match_template(img_dft.to(torch.complex64), tdft.to(torch.complex64), ctf, wf.float(), dfs,
                              torch.tensor([0.0]), eul, device=cuda_device, orientation_batch_size=8)
    z = got["scaled_mip"]
    for i, (Y, X) in enumerate(mg.centers_yx):
        y, x = int(round(Y)) - h // 2, int(round(X)) - h // 2
        win = z[y - 1:y + 2, x - 1:x + 2]
        assert float(win.max()) > 5.0
        iy, ix = np.unravel_index(int(win.argmax()), win.shape)
        assert abs(float(got["best_phi"][y - 1 + iy, x - 1 + ix]) - float(eul[i, 0])) < 1e-3


def test_match_peak_list_identical(cuda_device):
    """z-score peak list (Gaussian-noise cutoff, 5 px suppression) identical to the oracle's."""
    H = W = 512
    h = 64
    args = _inputs(H, W, h, 2, 1, 96, seed=11, n_particles=4)
    got = core_match_template(*args, device=cuda_device, orientation_batch_size=8)
    r64 = osr.core_match_template(*args, orientation_batch_size=8, dtype=torch.float64)
    cut = op.gaussian_noise_zscore_cutoff(H * W * 96 * 2, 1.0)
    z = r64["scaled_mip"]
    # the comparison needs a cutoff away from any oracle z (no pixel within 1e-3 of it)
    while bool(((z - cut).abs() < 1e-3).any()):
        cut += 2e-3
    gy, gx = op.find_peaks_from_zscore(got["scaled_mip"], cut)
    ry, rx = op.find_peaks_from_zscore(z.float(), cut)
    print(f"cutoff {cut:.3f}: {len(ry)} peaks", flush=True)
    assert len(ry) >= 4
    assert gy.tolist() == ry.tolist() and gx.tolist() == rx.tolist()
    for k in ("best_phi", "best_theta", "best_psi", "best_defocus"):
        assert got[k][gy, gx].tolist() == r64[k][ry, rx].float().tolist()


@pytest.mark.parametrize("opts", [
    {"colpass_columns": 1}, {"colpass_pair": 1}, {"rowpass_staging": 0}, {"rowpass_twiddles": 0},
    {"rowpass_twiddles": 0, "rowpass_staging": 0}, {"rowpass_rows": 4, "rowpass_twiddles": 0, "rowpass_staging": 1},
    {"rowpass_rows": 4, "rowpass_twiddles": 0, "rowpass_staging": 2}, {"rowpass_rows": 4}, {"colpass_chunks": 3}, {"colpass_columns": 1, "colpass_pair": 0},
    {"colpass_twiddles": 0}, {"colpass_twiddles": 0, "colpass_pair": 1}, {"colpass_columns": 1, "colpass_chunks": 5},
    {"rowpass_rows": 2}, {"rowpass_rows": 1}, {"rowpass_rows": 1, "rowpass_twiddles": 0, "rowpass_staging": 0},
    {"colpass_store": 1}, {"colpass_store": 1, "colpass_twiddles": 0},
    {"colpass_twiddles": 2}, {"colpass_twiddles": 2, "colpass_store": 1}, {"colpass_twiddles": 2, "colpass_pair": 1},
    {"colpass_twiddles": 2, "colpass_columns": 1}, {"colpass_twiddles": 1}, {"colpass_twiddles": 1, "colpass_store": 1},
    {"colpass_persistent": 0}, {"colpass_persistent": 0, "colpass_store": 0}, {"colpass_store": 0},
    {"colpass_persistent": 0, "colpass_twiddles": 0}, {"colpass_twiddles": 0, "colpass_store": 0},
    {"phase_timing": 1},
], ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()))
def test_launch_options_do_not_change_results(opts, cuda_device):
    from leopard_em_amd.backend import MatchPlan, raw_statistics

    args = _inputs(1024, 1024, 64, 2, 1, 8, seed=21)
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    out = []
    for o in ({}, opts):
        plan = MatchPlan(1024, 1024, 64, 4, 2, 1, cuda_device)
        for k, v in o.items():
            plan.set_option(k, v)
        plan.set_inputs(img_dft, ctf, wf)
        plan.reset_stats()
        plan.run(tdft.to(cuda_device), eul.to(cuda_device), 0, 8)
        torch.cuda.synchronize()
        out.append(raw_statistics(plan))
    for k in out[0]:
        assert torch.equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("opts", [{}, {"colpass_twiddles": 1}, {"colpass_store": 0}, {"colpass_persistent": 0},
                                  {"colpass_columns": 1}, {"colpass_columns": 1, "colpass_persistent": 0}],
                         ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()) or "defaults")
def test_launch_options_do_not_change_results_4096(opts, cuda_device):
    """The production image size (eight-wave four-step column transforms)."""
    from leopard_em_amd.backend import MatchPlan, raw_statistics

    print(f"options at 4096: {opts}", flush=True)
    H = W = 4096
    h, D = 64, 2
    gen = torch.Generator().manual_seed(5)
    img = torch.randn((H, W // 2 + 1), dtype=torch.complex64, generator=gen)
    img[0, 0] = 0
    tdft = torch.randn((h, h, h // 2 + 1), dtype=torch.complex64, generator=gen).to(cuda_device)
    ctf = torch.rand((1, D, h, h // 2 + 1), generator=gen)
    wf = torch.rand((h, h // 2 + 1), generator=gen)
    eul = (torch.rand((4, 3), generator=gen) * torch.tensor([360.0, 180.0, 360.0])).to(cuda_device)
    out = []
    for o in ({"colpass_twiddles": 0, "colpass_store": 0, "colpass_persistent": 0}, opts):
        plan = MatchPlan(H, W, h, 2, D, 1, cuda_device)
        for k, v in o.items():
            plan.set_option(k, v)
        plan.set_inputs(img, ctf, wf)
        plan.reset_stats()
        plan.run(tdft, eul, 0, 4)
        torch.cuda.synchronize()
        out.append(raw_statistics(plan))
        del plan
    for k in out[0]:
        assert torch.equal(out[0][k], out[1][k]), k


def test_unknown_launch_option_raises(cuda_device):
    from leopard_em_amd import _lib
    from leopard_em_amd.backend import MatchPlan

    plan = MatchPlan(256, 256, 64, 2, 1, 1, cuda_device)
    with pytest.raises(_lib.TdtmError, match="unknown option"):
        plan.set_option("no_such_option", 1)
    with pytest.raises(_lib.TdtmError):
        plan.set_option("colpass_columns", 3)
    plan.set_option("rowpass_staging", 0)
    plan.set_option("rowpass_twiddles", 1)
    plan.set_option("rowpass_rows", 4)
    with pytest.raises(_lib.TdtmError, match="LDS"):
        plan.set_option("rowpass_staging", 1)
    plan.set_option("rowpass_twiddles", 0)
    plan.set_option("rowpass_staging", 2)
    with pytest.raises(_lib.TdtmError, match="rowpass_staging"):
        plan.set_option("rowpass_twiddles", 1)
    with pytest.raises(_lib.TdtmError, match="rowpass_staging"):
        plan.set_option("rowpass_rows", 2)
    with pytest.raises(_lib.TdtmError, match="rowpass_rows"):
        plan.set_option("rowpass_rows", 3)


def test_phase_times(cuda_device):
    from leopard_em_amd.backend import MatchPlan

    args = _inputs(512, 512, 64, 2, 1, 8, seed=5)
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    plan = MatchPlan(512, 512, 64, 4, 2, 1, cuda_device)
    plan.set_inputs(img_dft, ctf, wf)
    plan.reset_stats()
    plan.set_option("phase_timing", 1)
    plan.run(tdft.to(cuda_device), eul.to(cuda_device), 0, 8)
    times, n = plan.phase_times()
    assert n == 2
    assert set(times) == {"project", "rowfwd", "colpass", "rowpass"}
    assert all(v > 0 for v in times.values())
    times, n = plan.phase_times()  # events were released
    assert n == 0 and all(v == 0 for v in times.values())


def test_async_staging_refused_for_sub_wave_row_blocks(cuda_device):
    """Row-pass blocks of W / 4 threads are below one wave for W < 256: the asynchronous
    staging modes are refused there (the synchronous default serves every size)."""
    from leopard_em_amd import _lib
    from leopard_em_amd.backend import MatchPlan

    plan = MatchPlan(128, 128, 32, 2, 1, 1, cuda_device)
    plan.set_option("rowpass_twiddles", 0)
    plan.set_option("rowpass_rows", 4)
    for mode in (1, 2):
        with pytest.raises(_lib.TdtmError, match="at least one wave"):
            plan.set_option("rowpass_staging", mode)
    plan.set_option("rowpass_staging", 0)


def test_image_size_change_recaptures_graph(cuda_device):
    """A captured step carries the image size of the projection normalisation: after
    tdtm_match_plan_set_image_size the next run must re-capture, matching a plan that had
    that size from the start (launched directly), and differing from the old size."""
    from leopard_em_amd import _lib
    from leopard_em_amd.backend import MatchPlan, raw_statistics

    print("image size change after a graph capture: 1024x1024 grid, 1000x900 image", flush=True)
    args = _inputs(1024, 1024, 64, 2, 1, 8, seed=43)
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    tdft_d, eul_d = tdft.to(cuda_device), eul.to(cuda_device)
    lib = _lib.load()

    g = MatchPlan(1024, 1024, 64, 4, 2, 1, cuda_device)
    g.set_inputs(img_dft, ctf, wf)
    g.reset_stats()
    g.run(tdft_d, eul_d, 0, 8)  # captured at 1024 x 1024
    torch.cuda.synchronize()
    before = raw_statistics(g)
    _lib.check(lib.tdtm_match_plan_set_image_size(g.handle, 1000, 900))
    g.reset_stats()
    g.run(tdft_d, eul_d, 0, 8)
    torch.cuda.synchronize()
    after = raw_statistics(g)

    d = MatchPlan(1024, 1024, 64, 4, 2, 1, cuda_device)
    d.set_option("graph", 0)
    _lib.check(lib.tdtm_match_plan_set_image_size(d.handle, 1000, 900))
    d.set_inputs(img_dft, ctf, wf)
    d.reset_stats()
    d.run(tdft_d, eul_d, 0, 8)
    torch.cuda.synchronize()
    ref = raw_statistics(d)
    for k in ref:
        assert torch.equal(ref[k], after[k]), k
    assert not torch.equal(before["mip"], after["mip"])


def test_magnification_and_option_changes_recapture_graph(cuda_device):
    """Captured steps carry the slice kernel's magnification and the launch shapes: after
    set_magnification or set_option the next run re-captures and matches a plan that had
    the setting from the start (launched directly)."""
    from leopard_em_amd.backend import MatchPlan, raw_statistics

    print("magnification / option change after a graph capture: 512x512 h64 D2 O8", flush=True)
    img_dft, tdft, ctf, wf, dfs, pix, eul = _inputs(512, 512, 64, 2, 1, 8, seed=47)
    tdft_d, eul_d = tdft.to(cuda_device), eul.to(cuda_device)
    mag = [[1.02, 0.01], [-0.015, 0.99]]

    def run(plan):
        plan.reset_stats()
        plan.run(tdft_d, eul_d, 0, 8)
        torch.cuda.synchronize()
        return raw_statistics(plan)

    def fresh(graph, **settings):
        p = MatchPlan(512, 512, 64, 4, 2, 1, cuda_device)
        p.set_option("graph", graph)
        if "mag" in settings:
            p.set_magnification(settings["mag"])
        for k, v in settings.get("opts", {}).items():
            p.set_option(k, v)
        p.set_inputs(img_dft, ctf, wf)
        return p

    g = fresh(1)
    plain = run(g)                      # captured without magnification
    g.set_magnification(mag)
    with_mag = run(g)
    ref = run(fresh(0, mag=mag))
    for k in ref:
        assert torch.equal(ref[k], with_mag[k]), k
    assert not torch.equal(plain["mip"], with_mag["mip"])
    g.set_option("rowpass_rows", 2)     # another launch shape, same results
    assert all(torch.equal(with_mag[k], v) for k, v in run(g).items())
    g.set_magnification(None)
    back = run(g)
    for k in plain:
        assert torch.equal(plain[k], back[k]), k


def test_graph_replay_bit_identical_to_direct_steps(cuda_device):
    """tdtm_match_run with graph replays (default) against the same batches launched as
    plain steps: every statistics map bit-identical.  The graph is replayed across two
    run() calls with eager work and a synchronisation between them, the last batch is
    partial (run as a plain step), and a second statistics set forces a re-capture."""
    from leopard_em_amd.backend import MatchPlan, raw_statistics

    print("graph replay vs direct steps: 1024x1024 h64 D2 O14 B4", flush=True)
    args = _inputs(1024, 1024, 64, 2, 1, 14, seed=41)
    img_dft, tdft, ctf, wf, dfs, pix, eul = args
    tdft_d, eul_d = tdft.to(cuda_device), eul.to(cuda_device)

    def fresh(graph):
        p = MatchPlan(1024, 1024, 64, 4, 2, 1, cuda_device)
        p.set_option("graph", graph)
        p.set_inputs(img_dft, ctf, wf)
        p.reset_stats()
        return p

    direct = fresh(0)
    for i in range(0, 14, 4):
        direct.step(tdft_d, eul_d, i, min(4, 14 - i))
    torch.cuda.synchronize()
    ref = raw_statistics(direct)

    g = fresh(1)
    g.run(tdft_d, eul_d, 0, 8)           # two replays
    junk = torch.randn(1 << 20, device=cuda_device).sum()  # eager work between replays
    torch.cuda.synchronize()
    g.run(tdft_d, eul_d, 8, 14)          # one replay + a partial batch of 2
    torch.cuda.synchronize()
    got = raw_statistics(g)
    for k in ref:
        assert torch.equal(ref[k], got[k]), k
    # new statistics tensors: the kept graph no longer matches and is re-captured
    g.mip = torch.empty_like(g.mip)
    g.best_index = torch.empty_like(g.best_index)
    g.sums = [torch.empty_like(s) for s in g.sums]
    g.reset_stats()
    g.run(tdft_d, eul_d, 0, 14)
    torch.cuda.synchronize()
    got2 = raw_statistics(g)
    for k in ref:
        assert torch.equal(ref[k], got2[k]), k
    assert float(junk) == float(junk)
