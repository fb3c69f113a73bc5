"""YAML-configured match-template runs: inputs from MRC files, maps and peaks out.

Counterpart of ``MatchTemplateManager`` and ``MatchTemplateResult``
(``pydantic_models/managers/match_template_manager.py:140-480``,
``pydantic_models/results/match_template_result.py:40-337``) for the configuration
subset the search engine runs:

* optics group (CTF parameters, anisotropic magnification matrix), relative defocus
  search, uniform orientation grid with C1 symmetry or explicit angle ranges, whitening,
  bandpass and arbitrary-curve filters (the phase-randomisation configuration is
  accepted and, as in the reference's combined filter, not applied:
  ``config/correlation_filters.py:395-396``);
* Zernike terms, other base grids and symmetries raise NotImplementedError when
  enabled;
* YAML keys, defaults and output names follow the reference's example configuration
  (``programs/match_template/match_template_example_config.yaml``); plain classes
  stand in for the pydantic models.

Additions: ``run_match_template(checkpoint_path=...)`` checkpoints and resumes the
search statistics (the reference lists this as TODO at
``match_template_result.py:51-56``).
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import Any, Optional

import pandas as pd
import torch
import yaml

from .analysis import (
    MatchTemplatePeaks,
    extract_peaks_and_statistics_zscore,
    match_template_peaks_to_dataframe,
    match_template_peaks_to_dict,
)
from .backend.core_match_template import core_match_template, core_match_template_resumable
from .io.mrc import load_mrc_image, load_mrc_volume, read_mrc_header, write_mrc_from_tensor
from .utils.preprocessing import ctf_filter_stack, prepare_image, volume_to_rfft_fourier_slice
from .utils.search import get_search_tensors, get_uniform_euler_angles


def _take(cfg: dict, cls_name: str, allowed: dict) -> dict:
    unknown = set(cfg) - set(allowed)
    if unknown:
        raise ValueError(f"{cls_name}: unknown field(s) {sorted(unknown)}")
    return {k: cfg.get(k, v) for k, v in allowed.items()}


class OpticsGroup:
    """``pydantic_models/data_structures/optics_group.py`` (CTF parameters)."""

    FIELDS = dict(label="", voltage=300.0, pixel_size=1.0, defocus_u=0.0, defocus_v=0.0, astigmatism_angle=0.0,
                  spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0, ctf_B_factor=0.0,
                  chromatic_aberration=0.0, mtf_reference=None, mtf_values=None, beam_tilt_x=None, beam_tilt_y=None,
                  odd_zernikes=None, even_zernikes=None, mag_matrix=None)

    def __init__(self, **cfg):
        for k, v in _take(cfg, "OpticsGroup", self.FIELDS).items():
            setattr(self, k, v)
        if self.odd_zernikes or self.even_zernikes:
            raise NotImplementedError("Zernike aberration terms are not supported by the tdtm backend")
        if self.mag_matrix is not None and len(self.mag_matrix) != 4:
            raise ValueError("mag_matrix must be a list of 4 floats")

    @property
    def mag_matrix_tensor(self) -> Optional[torch.Tensor]:
        """2 x 2 float32 tensor [[m0, m1], [m2, m3]] of ``mag_matrix`` or None
        (``optics_group.py:97-113``)."""
        if self.mag_matrix is None:
            return None
        return torch.tensor(self.mag_matrix, dtype=torch.float32).reshape(2, 2)


class DefocusSearchConfig:
    """``config/defocus_search.py:43-74``."""

    FIELDS = dict(enabled=True, defocus_min=-1000.0, defocus_max=1000.0, defocus_step=200.0,
                  skip_enforce_zero=False)

    def __init__(self, **cfg):
        for k, v in _take(cfg, "DefocusSearchConfig", self.FIELDS).items():
            setattr(self, k, v)

    @property
    def defocus_values(self) -> torch.Tensor:
        if not self.enabled:
            return torch.tensor([0.0])
        if self.defocus_max < self.defocus_min:
            raise ValueError("Defocus search parameters result in no values to search over!")
        return get_search_tensors(self.defocus_min, self.defocus_max, self.defocus_step, self.skip_enforce_zero)


class OrientationSearchConfig:
    """``config/orientation_search.py:127-170`` ('uniform' base grid)."""

    FIELDS = dict(psi_step=1.5, theta_step=2.5, phi_min=None, phi_max=None, theta_min=None, theta_max=None,
                  psi_min=None, psi_max=None, base_grid_method="uniform", symmetry="C1")

    def __init__(self, **cfg):
        for k, v in _take(cfg, "OrientationSearchConfig", self.FIELDS).items():
            setattr(self, k, v)
        if self.base_grid_method != "uniform":
            raise NotImplementedError(f"base grid {self.base_grid_method!r} is not supported (only 'uniform')")
        ranges = [self.phi_min, self.phi_max, self.theta_min, self.theta_max, self.psi_min, self.psi_max]
        if self.symmetry is not None and any(r is not None for r in ranges):
            raise ValueError("Please set all angle ranges to None when using symmetry.")
        if self.symmetry not in (None, "C1"):
            raise NotImplementedError(f"symmetry {self.symmetry!r} is not supported (only C1 or explicit ranges)")

    @property
    def euler_angles(self) -> torch.Tensor:
        def val(v, d):
            return d if v is None else v

        return get_uniform_euler_angles(
            psi_step=self.psi_step, theta_step=self.theta_step,
            phi_min=val(self.phi_min, 0.0), phi_max=val(self.phi_max, 360.0),
            theta_min=val(self.theta_min, 0.0), theta_max=val(self.theta_max, 180.0),
            psi_min=val(self.psi_min, 0.0), psi_max=val(self.psi_max, 360.0))


class PreprocessingFilters:
    """``config/correlation_filters.py``: whitening, bandpass and arbitrary-curve filters
    (phase randomisation is configured but left out of the combined filter, as in the
    reference, ``:395-396``)."""

    def __init__(self, **cfg):
        known = {"whitening_filter", "bandpass_filter", "phase_randomization_filter", "arbitrary_curve_filter"}
        unknown = set(cfg) - known
        if unknown:
            raise ValueError(f"PreprocessingFilters: unknown field(s) {sorted(unknown)}")
        wf = _take(cfg.get("whitening_filter") or {}, "WhiteningFilterConfig",
                   dict(enabled=True, num_freq_bins=None, max_freq=0.5, do_power_spectrum=True))
        self.whitening_enabled = bool(wf["enabled"])
        self.num_freq_bins = wf["num_freq_bins"]
        self.max_freq = 0.5 if wf["max_freq"] is None else float(wf["max_freq"])
        if not wf["do_power_spectrum"]:
            raise NotImplementedError("whitening from the amplitude spectrum (do_power_spectrum=False)")
        bp = _take(cfg.get("bandpass_filter") or {}, "BandpassFilterConfig",
                   dict(enabled=False, low_freq_cutoff=None, high_freq_cutoff=None, falloff=None))
        # None -> the reference's defaults (correlation_filters.py:256-259)
        self.bandpass = None
        if bp["enabled"]:
            self.bandpass = (0.0 if bp["low_freq_cutoff"] is None else float(bp["low_freq_cutoff"]),
                             1.0 if bp["high_freq_cutoff"] is None else float(bp["high_freq_cutoff"]),
                             0.0 if bp["falloff"] is None else float(bp["falloff"]))
        ac = _take(cfg.get("arbitrary_curve_filter") or {}, "ArbitraryCurveFilterConfig",
                   dict(enabled=False, frequencies=None, amplitudes=None))
        self.curve = None
        if ac["enabled"]:
            if ac["frequencies"] is None or ac["amplitudes"] is None:
                raise ValueError("When enabled, both 'frequencies' and 'amplitudes' must be provided.")
            self.curve = ([float(v) for v in ac["frequencies"]], [float(v) for v in ac["amplitudes"]])
        self.phase_randomization = _take(cfg.get("phase_randomization_filter") or {},
                                         "PhaseRandomizationFilterConfig", dict(enabled=False, cuton=None))
class ComputationalConfig:
    """``config/computational_config.py``: devices and the stream-count hint."""

    FIELDS = dict(gpu_ids=[0], num_cpus=1, backend="streamed")

    def __init__(self, **cfg):
        for k, v in _take(cfg, "ComputationalConfig", self.FIELDS).items():
            setattr(self, k, v)

    @property
    def gpu_devices(self) -> list[torch.device]:
        ids = self.gpu_ids
        if ids == "all":
            ids = list(range(torch.cuda.device_count()))
        if isinstance(ids, int):
            ids = [ids]
        return [torch.device(f"cuda:{i}") for i in ids]


# Peak-table columns, in the order of ``pydantic_models/formats.py:3-43`` (an
# interchange format read by the refine step).
MATCH_TEMPLATE_DF_COLUMN_ORDER = [
    "particle_index", "mip", "scaled_mip", "correlation_mean", "correlation_variance", "total_correlations",
    "pos_x", "pos_y", "pos_x_img", "pos_y_img", "pos_x_img_angstrom", "pos_y_img_angstrom", "phi", "theta", "psi",
    "relative_defocus", "defocus_u", "defocus_v", "astigmatism_angle", "pixel_size", "refined_pixel_size",
    "voltage", "spherical_aberration", "amplitude_contrast_ratio", "phase_shift", "ctf_B_factor",
    "even_zernikes", "odd_zernikes", "mag_matrix", "micrograph_path", "template_path", "mip_path",
    "scaled_mip_path", "psi_path", "theta_path", "phi_path", "defocus_path", "correlation_average_path",
    "correlation_variance_path",
]

_RESULT_PATHS = ["mip_path", "scaled_mip_path", "correlation_average_path", "correlation_variance_path",
                 "orientation_psi_path", "orientation_theta_path", "orientation_phi_path", "relative_defocus_path"]
_RESULT_MAPS = ["mip", "scaled_mip", "correlation_average", "correlation_variance", "orientation_psi",
                "orientation_theta", "orientation_phi", "relative_defocus"]


class MatchTemplateResult:
    """``results/match_template_result.py``: the eight maps, their paths and the peaks."""

    def __init__(self, allow_file_overwrite: bool = False, **paths):
        unknown = set(paths) - set(_RESULT_PATHS)
        if unknown:
            raise ValueError(f"MatchTemplateResult: unknown field(s) {sorted(unknown)}")
        self.allow_file_overwrite = allow_file_overwrite
        for k in _RESULT_PATHS:
            setattr(self, k, paths.get(k))
        for k in _RESULT_MAPS:
            setattr(self, k, None)
        self.total_projections = 0
        self.total_orientations = 0
        self.total_defocus = 0
        self.match_template_peaks: Optional[MatchTemplatePeaks] = None

    def apply_valid_cropping(self, template_shape: tuple[int, int]) -> None:
        """Crop every map to (H - h + 1, W - w + 1) (``match_template_result.py:206-236``)."""
        img_h, img_w = self.mip.shape
        h, w = template_shape
        for k in _RESULT_MAPS:
            setattr(self, k, getattr(self, k)[: img_h - h + 1, : img_w - w + 1])

    def load_tensors_from_paths(self) -> None:
        for m, p in zip(_RESULT_MAPS, _RESULT_PATHS):
            setattr(self, m, load_mrc_image(getattr(self, p)))

    def locate_peaks(self, **kwargs) -> MatchTemplatePeaks:
        self.match_template_peaks = extract_peaks_and_statistics_zscore(
            mip=self.mip, scaled_mip=self.scaled_mip, best_psi=self.orientation_psi,
            best_theta=self.orientation_theta, best_phi=self.orientation_phi, best_defocus=self.relative_defocus,
            correlation_average=self.correlation_average, correlation_variance=self.correlation_variance,
            total_correlation_positions=self.total_projections, **kwargs)
        return self.match_template_peaks

    def peaks_to_dict(self) -> dict:
        if self.match_template_peaks is None:
            self.locate_peaks()
        return match_template_peaks_to_dict(self.match_template_peaks)

    def peaks_to_dataframe(self) -> pd.DataFrame:
        if self.match_template_peaks is None:
            self.locate_peaks()
        return match_template_peaks_to_dataframe(self.match_template_peaks)

    def export_results(self) -> None:
        for m, p in zip(_RESULT_MAPS, _RESULT_PATHS):
            path = getattr(self, p)
            if path is None:
                raise ValueError(f"{p} is not set")
            write_mrc_from_tensor(getattr(self, m), path, overwrite=self.allow_file_overwrite)


class MatchTemplateManager:
    """Builds the search inputs from a configuration and runs the tdtm search."""

    def __init__(self, template_volume_path: str | None = None, micrograph_path: str | None = None,
                 optics_group: dict | None = None, defocus_search_config: dict | None = None,
                 orientation_search_config: dict | None = None, preprocessing_filters: dict | None = None,
                 computational_config: dict | None = None, match_template_result: dict | None = None,
                 pixel_size_search_config: dict | None = None,
                 micrograph: torch.Tensor | None = None, template_volume: torch.Tensor | None = None):
        if pixel_size_search_config and pixel_size_search_config.get("enabled", False):
            raise NotImplementedError("match template searches one pixel size (as the reference does)")
        self.template_volume_path = template_volume_path
        self.micrograph_path = micrograph_path
        self.optics_group = OpticsGroup(**(optics_group or {}))
        self.defocus_search_config = DefocusSearchConfig(**(defocus_search_config or {}))
        self.orientation_search_config = OrientationSearchConfig(**(orientation_search_config or {}))
        self.preprocessing_filters = PreprocessingFilters(**(preprocessing_filters or {}))
        self.computational_config = ComputationalConfig(**(computational_config or {}))
        self.match_template_result = MatchTemplateResult(**(match_template_result or {}))
        self.micrograph = micrograph
        self.template_volume = template_volume

    @classmethod
    def from_dict(cls, cfg: dict[str, Any]) -> "MatchTemplateManager":
        return cls(**cfg)

    @classmethod
    def from_yaml(cls, path: str | os.PathLike | Path) -> "MatchTemplateManager":
        with open(path, encoding="utf-8") as f:
            return cls.from_dict(yaml.safe_load(f))

    def to_dict(self) -> dict:
        og = {k: getattr(self.optics_group, k) for k in OpticsGroup.FIELDS if getattr(self.optics_group, k) is not None}
        return {
            "template_volume_path": self.template_volume_path,
            "micrograph_path": self.micrograph_path,
            "optics_group": og,
            "defocus_search_config": {k: getattr(self.defocus_search_config, k) for k in DefocusSearchConfig.FIELDS},
            "orientation_search_config": {k: getattr(self.orientation_search_config, k)
                                          for k in OrientationSearchConfig.FIELDS},
            "computational_config": {k: getattr(self.computational_config, k) for k in ComputationalConfig.FIELDS},
            "match_template_result": dict({k: getattr(self.match_template_result, k) for k in _RESULT_PATHS},
                                          allow_file_overwrite=self.match_template_result.allow_file_overwrite),
        }

    def to_yaml(self, path: str | os.PathLike | Path) -> None:
        with open(path, "w", encoding="utf-8") as f:
            yaml.safe_dump(self.to_dict(), f, sort_keys=False)

    def make_backend_core_function_kwargs(self) -> dict[str, Any]:
        """Inputs of ``core_match_template`` (``match_template_manager.py:140-218``)."""
        if self.micrograph is None:
            self.micrograph = load_mrc_image(self.micrograph_path)
        if self.template_volume is None:
            self.template_volume = load_mrc_volume(self.template_volume_path)
        image = torch.as_tensor(self.micrograph).to(torch.float32)
        template = torch.as_tensor(self.template_volume).to(torch.float32)
        defocus_values = self.defocus_search_config.defocus_values
        pixel_size_offsets = torch.tensor([0.0], dtype=torch.float32)
        og = self.optics_group
        pf = self.preprocessing_filters
        # with a GPU the inputs are built by the library's kernels on the first search
        # device; otherwise on the host with torch, as the reference does
        dev = self.computational_config.gpu_devices[0] if torch.cuda.is_available() else torch.device("cpu")
        image_pre, filt_tmp = prepare_image(image.to(dev), template.shape[-1], whitening=pf.whitening_enabled,
                                            num_freq_bins=pf.num_freq_bins, max_freq=pf.max_freq,
                                            bandpass=pf.bandpass, curve=pf.curve)
        ctf = ctf_filter_stack(
            (template.shape[0], template.shape[0]), og.defocus_u, og.defocus_v, og.astigmatism_angle,
            defocus_values, pixel_size_offsets, voltage=og.voltage, spherical_aberration=og.spherical_aberration,
            amplitude_contrast_ratio=og.amplitude_contrast_ratio, phase_shift=og.phase_shift,
            pixel_size=og.pixel_size, ctf_B_factor=og.ctf_B_factor, mag_matrix=og.mag_matrix_tensor, device=dev)
        return {
            "image_dft": image_pre.to(torch.complex64),
            "template_dft": volume_to_rfft_fourier_slice(template.to(dev)).to(torch.complex64),
            "ctf_filters": ctf,
            "whitening_filter_template": filt_tmp.to(torch.float32),
            "euler_angles": self.orientation_search_config.euler_angles.to(torch.float32),
            "defocus_values": defocus_values,
            "pixel_values": pixel_size_offsets,
            "device": self.computational_config.gpu_devices,
            "mag_matrix": og.mag_matrix_tensor,
        }

    def run_match_template(self, orientation_batch_size: int = 16, do_result_export: bool = True,
                           do_valid_cropping: bool = True, checkpoint_path: str | None = None,
                           checkpoint_interval_s: float = 600.0) -> bool:
        """Run the search and fill ``match_template_result``.

        With ``checkpoint_path`` the statistics are checkpointed every
        ``checkpoint_interval_s`` seconds and a matching checkpoint is resumed (one
        device).  Returns True when the search completed.
        """
        kw = self.make_backend_core_function_kwargs()
        if checkpoint_path is None:
            results = core_match_template(**kw, orientation_batch_size=orientation_batch_size,
                                          num_cuda_streams=self.computational_config.num_cpus,
                                          backend=self.computational_config.backend)
        else:
            devs = kw.pop("device")
            results = core_match_template_resumable(**kw, device=devs[0], checkpoint_path=checkpoint_path,
                                                    orientation_batch_size=orientation_batch_size,
                                                    checkpoint_interval_s=checkpoint_interval_s)
            if results is None:
                return False
        self._populate_match_template_result(results, do_result_export, do_valid_cropping)
        return True

    def _populate_match_template_result(self, results: dict, do_result_export: bool = True,
                                        do_valid_cropping: bool = True) -> None:
        r = self.match_template_result
        r.mip = results["mip"]
        r.scaled_mip = results["scaled_mip"]
        r.correlation_average = results["correlation_mean"]
        r.correlation_variance = results["correlation_variance"]
        r.orientation_psi = results["best_psi"]
        r.orientation_theta = results["best_theta"]
        r.orientation_phi = results["best_phi"]
        r.relative_defocus = results["best_defocus"]
        r.total_projections = results["total_projections"]
        r.total_orientations = results["total_orientations"]
        r.total_defocus = results["total_defocus"]
        r.match_template_peaks = None
        if do_valid_cropping:
            nx = self._template_size()
            r.apply_valid_cropping((nx, nx))
        if do_result_export:
            r.export_results()

    def _template_size(self) -> int:
        if self.template_volume is not None:
            return int(self.template_volume.shape[-1])
        return int(read_mrc_header(self.template_volume_path)["nx"])

    def results_to_dataframe(self, half_template_width_pos_shift: bool = True,
                             exclude_columns: Optional[list] = None,
                             locate_peaks_kwargs: Optional[dict] = None) -> pd.DataFrame:
        """Peak table with positions, poses, statistics and the optics group
        (``match_template_manager.py:360-480``)."""
        r = self.match_template_result
        if locate_peaks_kwargs is None:
            if r.match_template_peaks is None:
                r.locate_peaks()
        else:
            r.locate_peaks(**locate_peaks_kwargs)
        df = r.peaks_to_dataframe()
        nx = self._template_size()
        shift = nx // 2 if half_template_width_pos_shift else 0
        df["pos_y_img"] = df["pos_y"] + shift
        df["pos_x_img"] = df["pos_x"] + shift
        og = self.optics_group
        df["pos_y_img_angstrom"] = df["pos_y_img"] * og.pixel_size
        df["pos_x_img_angstrom"] = df["pos_x_img"] * og.pixel_size
        for k in ("defocus_u", "defocus_v", "astigmatism_angle", "pixel_size"):
            df[k] = getattr(og, k)
        df["refined_pixel_size"] = og.pixel_size
        for k in ("voltage", "spherical_aberration", "amplitude_contrast_ratio", "phase_shift", "ctf_B_factor"):
            df[k] = getattr(og, k)
        df["even_zernikes"] = [None] * len(df)
        df["odd_zernikes"] = [None] * len(df)
        df["mag_matrix"] = [og.mag_matrix] * len(df)
        df["micrograph_path"] = self.micrograph_path
        df["template_path"] = self.template_volume_path
        df["mip_path"] = r.mip_path
        df["scaled_mip_path"] = r.scaled_mip_path
        df["psi_path"] = r.orientation_psi_path
        df["theta_path"] = r.orientation_theta_path
        df["phi_path"] = r.orientation_phi_path
        df["defocus_path"] = r.relative_defocus_path
        df["correlation_average_path"] = r.correlation_average_path
        df["correlation_variance_path"] = r.correlation_variance_path
        df["particle_index"] = df.index
        df = df.reindex(columns=MATCH_TEMPLATE_DF_COLUMN_ORDER)
        if exclude_columns is not None:
            df = df.drop(columns=exclude_columns)
        return df
