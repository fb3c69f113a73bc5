"""MRC2014 I/O, search grids, z-score peak extraction and the manager harness (CPU)."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch
from scipy.stats import norm

from leopard_em_amd import analysis
from leopard_em_amd.io import mrc
from leopard_em_amd.manager import MATCH_TEMPLATE_DF_COLUMN_ORDER, MatchTemplateManager, MatchTemplateResult
from leopard_em_amd.utils import search
from oracle import geometry as ogeo
from oracle import peaks as opeaks

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
IMAGE = os.path.join(GOLDEN, "test_image.mrc")
VOLUME = os.path.join(GOLDEN, "test_volume.mrc")


# ----------------------------------------------------------------------------- MRC
def test_read_reference_fixtures():
    """Fixtures written by mrcfile for the reference's tests/utils/test_data_io.py."""
    img = mrc.read_mrc_to_numpy(IMAGE)
    vol = mrc.read_mrc_to_numpy(VOLUME)
    assert img.dtype == np.float32 and img.shape == (2, 2)
    assert vol.shape == (2, 2, 2)
    np.testing.assert_array_equal(img.ravel(), np.arange(4))
    np.testing.assert_array_equal(vol.ravel(), np.arange(8))
    t = mrc.read_mrc_to_tensor(IMAGE)
    assert isinstance(t, torch.Tensor) and t.ndim == 2


def test_load_image_and_volume_dimension_errors():
    assert mrc.load_mrc_image(IMAGE).ndim == 2
    assert mrc.load_mrc_volume(VOLUME).ndim == 3
    with pytest.raises(ValueError, match="MRC file is not two-dimensional"):
        mrc.load_mrc_image(VOLUME)
    with pytest.raises(ValueError, match="MRC file is not three-dimensional"):
        mrc.load_mrc_volume(IMAGE)


@pytest.mark.parametrize("fixture,data", [(IMAGE, np.arange(4, dtype=np.float32).reshape(2, 2)),
                                          (VOLUME, np.arange(8, dtype=np.float32).reshape(2, 2, 2))])
def test_writer_header_matches_mrcfile(tmp_path, fixture, data):
    """Every header word before the labels equals the one mrcfile wrote for the same data."""
    out = tmp_path / "x.mrc"
    mrc.write_mrc_from_numpy(data, out)
    a = np.frombuffer(open(fixture, "rb").read(), dtype="<i4")
    b = np.frombuffer(open(out, "rb").read(), dtype="<i4")
    assert a.size == b.size
    np.testing.assert_array_equal(a[:56], b[:56])
    np.testing.assert_array_equal(a[256:], b[256:])


@pytest.mark.parametrize("dtype", [np.float32, np.float16, np.int16, np.uint16, np.int8, np.complex64])
def test_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(0)
    data = (rng.standard_normal((3, 5, 7)) * 20).astype(dtype)
    if np.iscomplexobj(data):
        data = data + 1j * rng.standard_normal(data.shape).astype(np.float32)
    p = tmp_path / "r.mrc"
    mrc.write_mrc_from_numpy(data, p, voxel_size=1.06)
    back, hdr = mrc.read_mrc(p)
    assert back.dtype == data.dtype
    np.testing.assert_array_equal(back, data)
    assert hdr.voxel_size == pytest.approx((1.06, 1.06, 1.06), rel=1e-6)
    t = torch.from_numpy(rng.standard_normal((9, 4)).astype(np.float32))
    mrc.write_mrc_from_tensor(t, p, overwrite=True)
    assert torch.equal(mrc.read_mrc_to_tensor(p), t)


def test_big_endian_and_extended_header(tmp_path):
    data = np.arange(12, dtype=np.float32).reshape(3, 4)
    hdr = np.zeros(256, dtype=">i4")
    hdr[0:4] = (4, 3, 1, 2)
    hdr[16:19] = (1, 2, 3)
    hdr[23] = 80
    raw = bytearray(hdr.tobytes())
    raw[208:212] = b"MAP "
    raw[212:214] = b"\x11\x11"
    p = tmp_path / "be.mrc"
    with open(p, "wb") as f:
        f.write(raw)
        f.write(b"\x07" * 80)
        f.write(data.astype(">f4").tobytes())
    np.testing.assert_array_equal(mrc.read_mrc_to_numpy(p), data)


def test_write_refuses_overwrite_and_header(tmp_path):
    p = tmp_path / "o.mrc"
    mrc.write_mrc_from_numpy(np.zeros((2, 2), np.float32), p)
    with pytest.raises(ValueError, match="already exists"):
        mrc.write_mrc_from_numpy(np.zeros((2, 2), np.float32), p)
    with pytest.raises(NotImplementedError):
        mrc.write_mrc_from_numpy(np.zeros((2, 2), np.float32), p, mrc_header={}, overwrite=True)
    with pytest.raises(ValueError):
        mrc.write_mrc_from_numpy(np.zeros((2, 2), np.int64), p, overwrite=True)
    # float64 is stored as float32
    mrc.write_mrc_from_numpy(np.ones((2, 2)), p, overwrite=True)
    assert mrc.read_mrc_to_numpy(p).dtype == np.float32


def test_truncated_file_is_an_error(tmp_path):
    p = tmp_path / "t.mrc"
    mrc.write_mrc_from_numpy(np.zeros((8, 8), np.float32), p)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-4])
    with pytest.raises(ValueError, match="data block"):
        mrc.read_mrc(p)


# ----------------------------------------------------------------------------- grids
@pytest.mark.parametrize("psi,theta,ranges", [(90.0, 90.0, {}), (10.0, 5.0, {}), (7.5, 12.0, {}),
                                              (5.0, 4.0, dict(phi_max=120.0, theta_max=90.0, psi_max=180.0))])
def test_uniform_grid_matches_oracle(psi, theta, ranges):
    a = search.get_uniform_euler_angles(psi, theta, **ranges)
    b = ogeo.get_uniform_euler_angles(psi, theta, **ranges)
    assert torch.equal(a, b)


def test_uniform_grid_reference_count():
    """``tests/pydantic_models/test_orientation_search.py:106-110``: 24 orientations at 90/90."""
    assert search.get_uniform_euler_angles(90.0, 90.0).shape == (24, 3)


@pytest.mark.parametrize("args", [(-1200.0, 1200.0, 200.0), (-250.0, 350.0, 200.0), (100.0, 500.0, 100.0)])
def test_search_tensors_match_oracle(args):
    assert torch.equal(search.get_search_tensors(*args), ogeo.get_search_tensors(*args))


# ----------------------------------------------------------------------------- peaks
@pytest.mark.parametrize("n,fp", [(1e6, 1.0), (4096 * 4096 * 1584000 * 13, 1.0), (1e4, 0.5), (2, 1.0)])
def test_zscore_cutoff_is_normal_quantile(n, fp):
    """sqrt(2) erfcinv(2 fp / n) is the upper fp/n quantile of N(0, 1)."""
    z = analysis.gaussian_noise_zscore_cutoff(int(n), fp)
    assert z == pytest.approx(norm.isf(fp / int(n)), rel=1e-9, abs=1e-12)
    assert z == pytest.approx(opeaks.gaussian_noise_zscore_cutoff(int(n), fp), rel=1e-12, abs=1e-15)


def _brute_force_pick(vals, locs, r):
    order = np.argsort(-vals, kind="stable")
    picked = []
    for i in order:
        if all(np.hypot(*(locs[i] - locs[j])) >= r for j in picked):
            picked.append(i)
    return locs[picked]


@pytest.mark.parametrize("seed,r", [(0, 5.0), (1, 2.5), (2, 9.0), (3, 0.0), (4, 1.0)])
def test_peak_picking_matches_oracle_and_brute_force(seed, r):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((96, 128)).astype(np.float32)
    for _ in range(12):  # clustered high values so suppression matters
        y, x = rng.integers(0, 96), rng.integers(0, 128)
        z[max(0, y - 3):y + 4, max(0, x - 3):x + 4] += rng.uniform(3, 6)
    zt = torch.from_numpy(z)
    cutoff = 2.0
    py, px = analysis.find_peaks_from_zscore(zt, cutoff, r)
    oy, ox = opeaks.find_peaks_from_zscore(zt, cutoff, r)
    assert torch.equal(py, oy) and torch.equal(px, ox)
    locs = np.argwhere(z > cutoff)
    bf = _brute_force_pick(z[z > cutoff], locs, r)
    np.testing.assert_array_equal(np.stack([py.numpy(), px.numpy()], 1), bf)


def test_peak_picking_edge_cases():
    z = torch.zeros((8, 8))
    py, px = analysis.find_peaks_from_zscore(z, 1.0)
    assert py.numel() == 0 and px.numel() == 0
    z[3, 3] = 5.0
    z[3, 8 - 1] = 4.0   # distance 4 < 5: suppressed
    z[7, 0] = 3.0       # distance 5: kept (strict <)
    py, px = analysis.find_peaks_from_zscore(z, 1.0, 5.0)
    assert py.tolist() == [3, 7] and px.tolist() == [3, 0]
    out = analysis.filter_peaks_by_distance(torch.zeros(0), torch.zeros((0, 2), dtype=torch.long), 5.0)
    assert out.shape == (0, 2)


def test_extract_peaks_and_statistics():
    rng = np.random.default_rng(5)
    shape = (40, 50)
    maps = {k: torch.from_numpy(rng.standard_normal(shape).astype(np.float32)) for k in
            ["mip", "psi", "theta", "phi", "df", "mean", "var"]}
    z = torch.from_numpy(rng.standard_normal(shape).astype(np.float32))
    z[10, 10] = 9.0
    z[30, 40] = 8.0
    pk = analysis.extract_peaks_and_statistics_zscore(maps["mip"], z, maps["psi"], maps["theta"], maps["phi"],
                                                      maps["df"], maps["mean"], maps["var"], 1000)
    cut = opeaks.gaussian_noise_zscore_cutoff(2000 * 1000)
    oy, ox = opeaks.find_peaks_from_zscore(z, cut)
    assert torch.equal(pk.pos_y, oy) and torch.equal(pk.pos_x, ox)
    assert pk.pos_y.tolist() == [10, 30]
    assert torch.equal(pk.theta, maps["theta"][oy, ox])
    assert pk.total_correlations == 1000
    with pytest.warns(UserWarning, match="No peaks"):
        analysis.extract_peaks_and_statistics_zscore(maps["mip"], z, maps["psi"], maps["theta"], maps["phi"],
                                                     maps["df"], maps["mean"], maps["var"], 1000,
                                                     z_score_cutoff=100.0)


def test_valid_crop():
    m = torch.arange(8 * 10).reshape(8, 10)
    c = analysis.valid_crop(m, (3, 4))
    assert c.shape == (6, 7) and torch.equal(c, m[:6, :7])


# ----------------------------------------------------------------------------- manager
REF_YAML = {
    "template_volume_path": "t.mrc",
    "micrograph_path": "m.mrc",
    "match_template_result": {"allow_file_overwrite": True, **{k: f"out_{k}.mrc" for k in [
        "mip_path", "scaled_mip_path", "orientation_psi_path", "orientation_theta_path", "orientation_phi_path",
        "relative_defocus_path", "correlation_average_path", "correlation_variance_path"]}},
    "optics_group": {"label": "g", "voltage": 300.0, "pixel_size": 1.06, "defocus_u": 5200.0, "defocus_v": 4950.0,
                     "astigmatism_angle": 25.0, "spherical_aberration": 2.7, "amplitude_contrast_ratio": 0.07,
                     "phase_shift": 0.0, "ctf_B_factor": 60.0},
    "defocus_search_config": {"defocus_max": 1200.0, "defocus_min": -1200.0, "defocus_step": 200.0,
                              "enabled": True},
    "orientation_search_config": {"base_grid_method": "uniform", "psi_step": 1.5, "theta_step": 2.5,
                                  "symmetry": "C1"},
    "preprocessing_filters": {"whitening_filter": {"enabled": True, "do_power_spectrum": True, "max_freq": 0.5,
                                                   "num_freq_bins": None},
                              "bandpass_filter": {"enabled": False, "falloff": None, "high_freq_cutoff": None,
                                                  "low_freq_cutoff": None}},
    "computational_config": {"gpu_ids": [0], "num_cpus": 1},
}


def test_manager_reads_reference_style_yaml(tmp_path):
    import yaml

    p = tmp_path / "cfg.yaml"
    p.write_text(yaml.safe_dump(REF_YAML))
    m = MatchTemplateManager.from_yaml(p)
    assert m.defocus_search_config.defocus_values.tolist() == [float(v) for v in range(-1200, 1201, 200)]
    # the production grid of the reference's benchmark (2.5 / 1.5 degrees)
    assert m.orientation_search_config.euler_angles.shape == (1584000, 3)
    m.to_yaml(tmp_path / "back.yaml")
    m2 = MatchTemplateManager.from_yaml(tmp_path / "back.yaml")
    assert m2.optics_group.defocus_u == 5200.0
    bad = dict(REF_YAML, orientation_search_config={"psi_step": 1.5, "theta_step": 2.5, "symmetry": "C2"})
    with pytest.raises(NotImplementedError):
        MatchTemplateManager.from_dict(bad)
    bad = dict(REF_YAML, defocus_search_config={"defocus_max": 1.0, "nonsense": 2})
    with pytest.raises(ValueError, match="unknown"):
        MatchTemplateManager.from_dict(bad)


def test_result_export_crop_and_dataframe(tmp_path):
    rng = np.random.default_rng(9)
    H, W, h = 48, 64, 16
    cfg = dict(REF_YAML)
    cfg["match_template_result"] = dict(REF_YAML["match_template_result"])
    for k in list(cfg["match_template_result"]):
        if k.endswith("_path"):
            cfg["match_template_result"][k] = str(tmp_path / cfg["match_template_result"][k])
    m = MatchTemplateManager.from_dict(cfg)
    m.template_volume = torch.zeros((h, h, h))
    z = torch.from_numpy(rng.standard_normal((H, W)).astype(np.float32))
    z[5, 7] = 12.0
    z[40, 60] = 12.0  # outside the valid region (H-h+1, W-w+1) = (33, 49)
    res = {"mip": z * 0.1, "scaled_mip": z, "best_phi": z + 1, "best_theta": z + 2, "best_psi": z + 3,
           "best_defocus": z + 4, "correlation_mean": z + 5, "correlation_variance": z + 6,
           "total_projections": 500, "total_orientations": 100, "total_defocus": 5}
    m._populate_match_template_result(res, do_result_export=True, do_valid_cropping=True)
    r = m.match_template_result
    assert r.mip.shape == (H - h + 1, W - h + 1)
    df = m.results_to_dataframe()
    assert list(df.columns) == MATCH_TEMPLATE_DF_COLUMN_ORDER
    assert df["pos_y"].tolist() == [5] and df["pos_x"].tolist() == [7]
    assert df["pos_y_img"].tolist() == [5 + h // 2]
    assert df["theta"].tolist() == [pytest.approx(14.0)]
    r2 = MatchTemplateResult(**{k: getattr(r, k) for k in ["mip_path", "scaled_mip_path", "correlation_average_path",
                                                           "correlation_variance_path", "orientation_psi_path",
                                                           "orientation_theta_path", "orientation_phi_path",
                                                           "relative_defocus_path"]})
    r2.load_tensors_from_paths()
    assert torch.equal(r2.scaled_mip, r.scaled_mip) and torch.equal(r2.relative_defocus, r.relative_defocus)


# ----------------------------------------------------------------------------- product preprocessing
def test_product_preprocessing_matches_oracle():
    from leopard_em_amd import synthetic
    from leopard_em_amd.utils import preprocessing as pp
    from oracle import filters as of

    h = 32
    t = synthetic.make_blob_template(4, h, 100, 10.0)
    mg = synthetic.make_micrograph(t, 4, (128, 192), 2, 12200.0, 11800.0, 30.0, snr=0.5)
    pre, wf = pp.prepare_image(mg.image, h)
    opre, owf = of.prepare_image_dft(mg.image, (h, h // 2 + 1))
    assert float((pre - opre).abs().max() / opre.abs().max()) < 1e-5
    assert float((wf - owf.float()).abs().max()) < 1e-6
    dfs = torch.tensor([-400.0, 0.0, 400.0])
    pix = torch.tensor([-0.02, 0.0, 0.03])
    ctf = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, pix, voltage=300.0, spherical_aberration=2.7,
                              amplitude_contrast_ratio=0.07, phase_shift=10.0, pixel_size=1.06, ctf_B_factor=60.0)
    octf = of.calculate_ctf_filter_stack_full_args(
        (h, h), 12200.0, 11800.0, dfs, pix, astigmatism_angle=30.0, voltage=300.0, spherical_aberration=2.7,
        amplitude_contrast_ratio=0.07, phase_shift=10.0, pixel_size=1.06, ctf_B_factor=60.0)
    assert ctf.shape == octf.shape == (3, 3, h, h // 2 + 1)
    assert float((ctf.double() - octf).abs().max()) < 1e-6
    vol = torch.from_numpy(np.random.default_rng(1).standard_normal((h, h, h)).astype(np.float32))
    assert torch.allclose(pp.volume_to_rfft_fourier_slice(vol), of.volume_to_rfft_fourier_slice(vol))


def test_manager_filters_and_magnification_on_the_host():
    """Bandpass / curve filters and the magnification matrix from the configuration reach
    the search inputs (host path: CPU tensors), matching the oracle's manager chain
    (match_template_manager.py:140-218); phase randomisation is accepted and not applied,
    as in the reference's combined filter (correlation_filters.py:395-396)."""
    from leopard_em_amd import synthetic
    from oracle import filters as of

    t = synthetic.make_blob_template(2, 32, 40, 10.0)
    mg = synthetic.make_micrograph(t, 3, (128, 128), 2, 5200.0, 4950.0, 25.0, snr=0.5)
    cfg = dict(REF_YAML)
    cfg["optics_group"] = dict(REF_YAML["optics_group"], mag_matrix=[1.01, 0.02, -0.01, 0.99])
    cfg["defocus_search_config"] = {"defocus_max": 200.0, "defocus_min": -200.0, "defocus_step": 200.0}
    cfg["preprocessing_filters"] = {
        "whitening_filter": {"enabled": True},
        "bandpass_filter": {"enabled": True, "low_freq_cutoff": 0.02, "high_freq_cutoff": 0.4, "falloff": 0.05},
        "arbitrary_curve_filter": {"enabled": True, "frequencies": [0.1, 0.3], "amplitudes": [1.0, 0.5]},
        "phase_randomization_filter": {"enabled": True, "cuton": 0.4},
    }
    m = MatchTemplateManager.from_dict(dict(cfg, micrograph=mg.image, template_volume=t.volume()))
    m.computational_config.gpu_ids = []  # host path even where a GPU exists
    import torch as _t

    if _t.cuda.is_available():
        pytest.skip("host-path check (the manager builds inputs on the GPU when one is present)")
    kw = m.make_backend_core_function_kwargs()
    want, wf = of.prepare_image_dft(mg.image.double(), (32, 17), bandpass=(0.02, 0.4, 0.05),
                                    curve=([0.1, 0.3], [1.0, 0.5]))
    rel = float((kw["image_dft"].to(_t.complex128) - want).abs().max() / want.abs().max())
    assert rel <= 1e-5
    assert float((kw["whitening_filter_template"].double() - wf).abs().max()) <= 1e-6
    mag = _t.tensor([[1.01, 0.02], [-0.01, 0.99]])
    assert _t.equal(kw["mag_matrix"], mag)
    ctf = of.calculate_ctf_filter_stack_full_args(
        (32, 32), 5200.0, 4950.0, _t.tensor([-200.0, 0.0, 200.0]), _t.tensor([0.0]), astigmatism_angle=25.0,
        voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0, pixel_size=1.06,
        ctf_B_factor=60.0, mag_matrix=mag.double())
    assert float((kw["ctf_filters"].double() - ctf).abs().max()) <= 1e-6
    bad = dict(cfg, preprocessing_filters={"arbitrary_curve_filter": {"enabled": True, "frequencies": [0.1]}})
    with pytest.raises(ValueError, match="amplitudes"):
        MatchTemplateManager.from_dict(bad)
