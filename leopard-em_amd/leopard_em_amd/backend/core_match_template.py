"""Whole-orientation search on the gfx950 HIP path.

Drop-in for ``leopard_em.backend.core_match_template.core_match_template``
(``src/leopard_em/backend/core_match_template.py:146-363``): same signature, same
output dictionary.  Differences, all deliberate:

* one process; every device in ``device`` gets a contiguous share of the
  orientations and the partial statistics are merged on the first device with the
  reference's rule (max, first device on ties; sums added) instead of the
  multiprocessing work queue (``backend/distributed.py``).
* sums and sums of squares are compensated (hi, lo) float32 pairs and the final
  mean / std / z-score are computed in float64 on the device.
* the best index of a partial last orientation batch is exact (the reference's
  ``arange(i, i + batch)`` at ``core_match_template.py:513-518`` mis-maps it when
  num_defocus * num_pixel > 1).
* ``num_cuda_streams`` and ``backend`` are accepted for signature compatibility; all
  work of a step is one fused launch sequence on the current stream.
* with the environment variable ``TDTM_AUTOTUNE=1`` each device's plan picks its
  launch shapes with ``backend.autotune`` (cached per shape and device).
* ``TDTM_FFT_ENGINE=1`` (fp16) or ``=2`` (bf16) runs the 4096-point column transforms of
  4096-row search grids on the matrix cores (radix-16 DFT matrix products with hi / lo
  split matrices and FP32 accumulation; results within 1e-3 of the FP64 reference
  instead of the FP32 path's 1e-4); other grids keep the FP32 path.
* image sizes that are not powers of two (e.g. 4092 x 5760 K3 micrographs) run in the
  next power-of-two grid with the image zero-padded at the bottom / right: every
  position whose template footprint lies inside the image (the valid region the peak
  search uses) gets the reference's value; the positions near the bottom / right edges
  whose footprint crosses the edge see zero padding instead of the reference's
  periodic wrap-around.
"""

from __future__ import annotations

import json
import os
import time
import warnings
import zlib

import numpy as np
import torch

from .. import _lib

CHECKPOINT_FORMAT = "tdtm-match-stats-1"

MAX_GRID_H, MAX_GRID_W = 4096, 8192


def _pow2(n: int) -> bool:
    return n > 0 and n & (n - 1) == 0


def search_grid(image_h: int, image_w: int) -> tuple[int, int]:
    """FFT grid of the search for an image size: the size itself when both edges are
    powers of two the kernels support, else the next powers of two (at least 64)."""
    if _pow2(image_h) and _pow2(image_w) and 64 <= image_h <= MAX_GRID_H and 64 <= image_w <= MAX_GRID_W:
        return image_h, image_w

    def up(n: int) -> int:
        g = 64
        while g < n:
            g *= 2
        return g

    gh, gw = up(image_h), up(image_w)
    if gh > MAX_GRID_H or gw > MAX_GRID_W:
        raise _lib.TdtmError(f"image {image_h}x{image_w} exceeds the largest search grid "
                             f"{MAX_GRID_H}x{MAX_GRID_W} (height x width)")
    return gh, gw


class MatchPlan:
    """Device plan + statistics maps for one search on one device.

    ``H`` x ``W`` is the image; the plan runs on the FFT grid ``Hg`` x ``Wg``
    (``search_grid``), and the statistics maps have the grid's shape (the image is their
    top-left ``H`` x ``W`` block)."""

    def __init__(self, image_h: int, image_w: int, template_h: int, orientation_batch_size: int,
                 num_defocus: int, num_pixel: int, device: torch.device):
        if device.type != "cuda":
            raise _lib.TdtmError(f"the tdtm backend runs on a HIP device, got {device}")
        self.lib = _lib.load()
        self.device = device
        self.H, self.W, self.h = image_h, image_w, template_h
        self.Hg, self.Wg = search_grid(image_h, image_w)
        self.padded = (self.Hg, self.Wg) != (image_h, image_w)
        self.B, self.D, self.C = orientation_batch_size, num_defocus, num_pixel
        handle = _lib.c_void_p()
        with torch.cuda.device(device):
            _lib.check(self.lib.tdtm_match_plan_create(self.Hg, self.Wg, template_h, orientation_batch_size,
                                                       num_defocus, num_pixel, _lib.ctypes.byref(handle)))
        self.handle = handle
        if self.padded:
            _lib.check(self.lib.tdtm_match_plan_set_image_size(handle, image_h, image_w))
        n = (self.Hg, self.Wg)
        self.mip = torch.empty(n, dtype=torch.float32, device=device)
        self.best_index = torch.empty(n, dtype=torch.int32, device=device)
        self.sums = [torch.empty(n, dtype=torch.float32, device=device) for _ in range(4)]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.tdtm_match_plan_destroy(h)
            except Exception:  # noqa: BLE001 -- interpreter shutdown
                pass
            self.handle = None

    @property
    def bytes(self) -> int:
        return int(self.lib.tdtm_match_plan_bytes(self.handle))

    def set_option(self, key: str, value: int) -> None:
        """Launch-shape option (see tdtm_match_plan_set_option); results are unaffected."""
        _lib.check(self.lib.tdtm_match_plan_set_option(self.handle, key.encode(), int(value)))

    def set_magnification(self, mag_matrix) -> None:
        """Anisotropic magnification matrix (2 x 2, or a flat list of 4 as in
        ``OpticsGroup.mag_matrix``; None = off) of the slices and the refinement CTF
        (``tdtm_match_plan_set_magnification``; ``cross_correlation.py:81-88``)."""
        _lib.check(self.lib.tdtm_match_plan_set_magnification(self.handle, mag_array(mag_matrix)))

    PHASES = ("project", "rowfwd", "colpass", "rowpass")

    def phase_times(self) -> tuple[dict[str, float], int]:
        """({phase: device ms summed over steps}, steps) for the steps recorded since the
        last call; needs ``set_option("phase_timing", 1)``."""
        out = (_lib.ctypes.c_double * len(self.PHASES))()
        n = _lib.c_int()
        _lib.check(self.lib.tdtm_match_plan_phase_times(self.handle, out, _lib.ctypes.byref(n)))
        return dict(zip(self.PHASES, list(out))), int(n.value)

    def _stream(self):
        return _lib.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_inputs(self, image_dft: torch.Tensor, ctf_filters: torch.Tensor, whitening: torch.Tensor) -> None:
        st = self._stream()
        if tuple(image_dft.shape) != (self.H, self.W // 2 + 1):
            raise ValueError(f"image_dft shape {tuple(image_dft.shape)} != {(self.H, self.W // 2 + 1)}")
        image = image_dft.to(self.device, torch.complex64).contiguous()
        if self.padded:
            image = self.regrid(image, st)
        self._image = image
        self._ctf = ctf_filters.to(self.device, torch.float32).contiguous()
        self._wf = whitening.to(self.device, torch.float32).contiguous()
        _lib.check(self.lib.tdtm_match_set_image(self.handle, _lib.ptr(self._image), st))
        _lib.check(self.lib.tdtm_match_set_filters(self.handle, _lib.ptr(self._ctf), _lib.ptr(self._wf), st))

    def regrid(self, image: torch.Tensor, st) -> torch.Tensor:
        """The spectrum of the image zero-padded into the grid: any-size inverse DFT
        into a zeroed grid-size array, then the grid's rfft2 (both on the device)."""
        with torch.cuda.device(self.device):
            real = torch.zeros((self.Hg, self.Wg), dtype=torch.float32, device=self.device)
            nbytes = self.lib.tdtm_irfft2_any_workspace_bytes(self.H, self.W)
            if nbytes < 0:
                raise _lib.TdtmError(f"image {self.H}x{self.W}: unsupported size for re-gridding")
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            _lib.check(self.lib.tdtm_irfft2_any(_lib.ptr(image), self.H, self.W, _lib.ptr(real), self.Wg,
                                                _lib.ptr(ws), st))
            out = torch.empty((self.Hg, self.Wg // 2 + 1), dtype=torch.complex64, device=self.device)
            _lib.check(self.lib.tdtm_rfft2(_lib.ptr(real), self.Hg, self.Wg, _lib.ptr(out), st))
        return out

    def image_view(self, t: torch.Tensor) -> torch.Tensor:
        """The image's H x W block of a grid-size map."""
        return t[:self.H, :self.W] if self.padded else t

    def reset_stats(self) -> None:
        _lib.check(self.lib.tdtm_stats_init(self.Hg * self.Wg, _lib.ptr(self.mip), _lib.ptr(self.best_index),
                                            *[_lib.ptr(s) for s in self.sums], self._stream()))

    def step(self, template_dft: torch.Tensor, euler_deg: torch.Tensor, start: int, count: int) -> None:
        _lib.check(self.lib.tdtm_match_step(
            self.handle, _lib.ptr(template_dft), _lib.ptr(euler_deg), euler_deg.shape[0], start, count,
            _lib.ptr(self.mip), _lib.ptr(self.best_index), *[_lib.ptr(s) for s in self.sums], self._stream()))

    def run(self, template_dft: torch.Tensor, euler_deg: torch.Tensor, start: int, end: int) -> None:
        """Orientations [start, end) in batches of B (tdtm_match_run: full batches replayed
        from a captured HIP graph unless the plan option "graph" is 0)."""
        _lib.check(self.lib.tdtm_match_run(
            self.handle, _lib.ptr(template_dft), _lib.ptr(euler_deg), euler_deg.shape[0], start, end,
            _lib.ptr(self.mip), _lib.ptr(self.best_index), *[_lib.ptr(s) for s in self.sums], self._stream()))

    def merge_from(self, other: "MatchPlan") -> None:
        """self.stats <- merge(self.stats, other.stats); other may live on another device."""
        tens = [other.mip, other.best_index, *other.sums]
        if other.device != self.device:
            tens = [t.to(self.device) for t in tens]
        _lib.check(self.lib.tdtm_stats_merge(
            self.Hg * self.Wg, _lib.ptr(self.mip), _lib.ptr(self.best_index), *[_lib.ptr(s) for s in self.sums],
            *[_lib.ptr(t) for t in tens], self._stream()))
        self._keep = tens

    def save_state(self, path: str, next_orientation: int, fingerprint: dict) -> None:
        """Write the running statistics and the next orientation index to ``path``.

        The maps are exact copies (mip, best index, compensated sums), so a search
        resumed from the file reproduces an uninterrupted run bit for bit.  The file is
        a numpy ``.npz`` written to a temporary name and renamed into place.
        """
        meta = dict(fingerprint, next_orientation=int(next_orientation), format=CHECKPOINT_FORMAT)
        arrays = {"mip": self.mip, "best_index": self.best_index, "sum_hi": self.sums[0],
                  "sum_lo": self.sums[1], "sq_hi": self.sums[2], "sq_lo": self.sums[3]}
        host = {k: v.cpu().numpy() for k, v in arrays.items()}
        host["meta"] = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), dtype=np.uint8)
        tmp = f"{path}.tmp{os.getpid()}.npz"
        np.savez(tmp, **host)
        os.replace(tmp, path)

    def load_state(self, path: str, fingerprint: dict) -> int:
        """Restore statistics saved by :meth:`save_state`; returns the next orientation.

        Raises ValueError when the file belongs to a different search (shapes, grid,
        defocus values or batch size differ).
        """
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            if meta.get("format") != CHECKPOINT_FORMAT:
                raise ValueError(f"{path}: unknown checkpoint format {meta.get('format')!r}")
            for k, v in fingerprint.items():
                if meta.get(k) != v:
                    raise ValueError(f"{path}: checkpoint is for a different search ({k}: {meta.get(k)!r} != {v!r})")
            dst = [self.mip, self.best_index, *self.sums]
            for name, t in zip(["mip", "best_index", "sum_hi", "sum_lo", "sq_hi", "sq_lo"], dst):
                a = z[name]
                if tuple(a.shape) != tuple(t.shape):
                    raise ValueError(f"{path}: map {name} has shape {a.shape}, expected {tuple(t.shape)}")
                t.copy_(torch.from_numpy(a))
        return int(meta["next_orientation"])

    def finalize(self, euler_deg: torch.Tensor, defocus_values: torch.Tensor, total: int) -> dict:
        """Output maps of the image (H x W): mip plus the decoded / scaled maps."""
        dev = self.device
        maps = [self.mip, self.best_index, *self.sums]
        if self.padded:
            maps = [self.image_view(m).contiguous() for m in maps]
        outs = [torch.empty((self.H, self.W), dtype=torch.float32, device=dev) for _ in range(7)]
        _lib.check(self.lib.tdtm_match_finalize(
            self.H * self.W, total, _lib.ptr(maps[0]), _lib.ptr(maps[1]),
            *[_lib.ptr(s) for s in maps[2:]], _lib.ptr(euler_deg), euler_deg.shape[0],
            _lib.ptr(defocus_values), defocus_values.shape[0], *[_lib.ptr(o) for o in outs], self._stream()))
        keys = ["scaled_mip", "correlation_mean", "correlation_variance", "best_phi", "best_theta", "best_psi",
                "best_defocus"]
        return dict(zip(keys, outs), mip=maps[0])


def mag_array(mag_matrix):
    """Host doubles {a00, a01, a10, a11} of a 2 x 2 magnification matrix (tensor, nested or
    flat list of 4, ``OpticsGroup.mag_matrix_tensor`` order) or None."""
    if mag_matrix is None:
        return None
    m = torch.as_tensor(mag_matrix, dtype=torch.float64).detach().cpu().reshape(-1)
    if m.numel() != 4:
        raise ValueError(f"mag_matrix must have 4 entries (2 x 2), got {m.numel()}")
    return (_lib.ctypes.c_double * 4)(*[float(v) for v in m])


def _as_device_list(device) -> list[torch.device]:
    if isinstance(device, (list, tuple)):
        devs = [torch.device(d) for d in device]
    else:
        devs = [torch.device(device)]
    if not devs:
        raise ValueError("no device given")
    return devs


# pylint: disable=too-many-arguments,too-many-locals
def core_match_template(
    image_dft: torch.Tensor,
    template_dft: torch.Tensor,
    ctf_filters: torch.Tensor,
    whitening_filter_template: torch.Tensor,
    defocus_values: torch.Tensor,
    pixel_values: torch.Tensor,
    euler_angles: torch.Tensor,
    device: torch.device | list[torch.device],
    orientation_batch_size: int = 1,
    num_cuda_streams: int = 1,
    backend: str = "streamed",
    mag_matrix: torch.Tensor | None = None,
) -> dict[str, torch.Tensor]:
    """Whole-orientation 2DTM search; signature and outputs of the reference's
    ``core_match_template`` (``backend/core_match_template.py:146-363``).

    Returns CPU tensors: mip, scaled_mip, best_phi, best_theta, best_psi,
    best_defocus, correlation_mean, correlation_variance (the standard deviation, as
    in the reference), and the ints total_projections, total_orientations,
    total_defocus.
    """
    _ = backend
    total_cc_per_batch = orientation_batch_size * defocus_values.shape[0] * pixel_values.shape[0]
    if num_cuda_streams > total_cc_per_batch:
        warnings.warn(
            f"Number of CUDA streams ({num_cuda_streams}) is greater than the number of "
            f"cross-correlations per batch ({total_cc_per_batch}).", stacklevel=2)
    devs = _as_device_list(device)
    H = image_dft.shape[0]
    W = 2 * (image_dft.shape[1] - 1)
    h = template_dft.shape[1]
    O, D, C = euler_angles.shape[0], defocus_values.shape[0], pixel_values.shape[0]
    if ctf_filters.shape != (C, D, h, h // 2 + 1):
        raise ValueError(f"ctf_filters shape {tuple(ctf_filters.shape)} != {(C, D, h, h // 2 + 1)}")
    if whitening_filter_template.shape != (h, h // 2 + 1):
        raise ValueError("whitening_filter_template must have shape (h, h//2+1)")
    if template_dft.shape != (h, h, h // 2 + 1):
        raise ValueError("template_dft must have shape (h, h, h//2+1)")
    batch = max(1, min(orientation_batch_size, O))

    plans = []
    share = (O + len(devs) - 1) // len(devs)
    for r, dev in enumerate(devs):
        lo, hi = r * share, min(O, (r + 1) * share)
        plan = MatchPlan(H, W, h, batch, D, C, dev)
        if mag_matrix is not None:
            plan.set_magnification(mag_matrix)
        engine = int(os.environ.get("TDTM_FFT_ENGINE", "0"))
        if engine and plan.Hg == 4096:
            plan.set_option("fft_engine", engine)
        with torch.cuda.device(dev):
            plan.set_inputs(image_dft, ctf_filters, whitening_filter_template)
            tdft = template_dft.to(dev, torch.complex64).contiguous()
            eul = euler_angles.to(dev, torch.float32).contiguous()
            plan.reset_stats()
            if os.environ.get("TDTM_AUTOTUNE", "0") == "1" and hi > lo:
                from .autotune import autotune

                autotune(plan, tdft, eul[lo:hi].contiguous())
            if hi > lo:
                plan.run(tdft, eul, lo, hi)
            plan._tdft, plan._eul = tdft, eul
        plans.append(plan)
    main = plans[0]
    with torch.cuda.device(main.device):
        for other in plans[1:]:
            torch.cuda.synchronize(other.device)
            main.merge_from(other)
    return _finish(main, defocus_values, O, D, C)


def _finish(main: MatchPlan, defocus_values: torch.Tensor, O: int, D: int, C: int) -> dict:
    with torch.cuda.device(main.device):
        dvals = defocus_values.to(main.device, torch.float32).contiguous()
        total = O * D * C
        maps = main.finalize(main._eul, dvals, total)
        torch.cuda.synchronize(main.device)
    return {
        "mip": maps["mip"].cpu(),
        "scaled_mip": maps["scaled_mip"].cpu(),
        "best_phi": maps["best_phi"].cpu(),
        "best_theta": maps["best_theta"].cpu(),
        "best_psi": maps["best_psi"].cpu(),
        "best_defocus": maps["best_defocus"].cpu(),
        "correlation_mean": maps["correlation_mean"].cpu(),
        "correlation_variance": maps["correlation_variance"].cpu(),
        "total_projections": total,
        "total_orientations": O,
        "total_defocus": D,
    }


def _crc(t: torch.Tensor) -> int:
    a = t.detach().contiguous().cpu()
    if a.is_complex():
        a = torch.view_as_real(a)
    return zlib.crc32(a.numpy().tobytes())


def search_fingerprint(image_dft, template_dft, ctf_filters, whitening_filter_template, defocus_values,
                       pixel_values, euler_angles, orientation_batch_size, mag_matrix=None) -> dict:
    """Identity of a search for checkpoint matching: shapes, batch size, magnification and
    CRC-32s of every input."""
    fp = {
        "image_shape": list(image_dft.shape),
        "template_shape": list(template_dft.shape),
        "num_orientations": int(euler_angles.shape[0]),
        "num_defocus": int(defocus_values.shape[0]),
        "num_pixel": int(pixel_values.shape[0]),
        "orientation_batch_size": int(orientation_batch_size),
        "crc": [_crc(t) for t in (image_dft, template_dft, ctf_filters, whitening_filter_template,
                                  defocus_values, pixel_values, euler_angles)],
    }
    if mag_matrix is not None:
        fp["mag_matrix"] = [float(v) for v in torch.as_tensor(mag_matrix, dtype=torch.float64).reshape(-1)]
    return fp


# pylint: disable=too-many-arguments,too-many-locals
def core_match_template_resumable(
    image_dft: torch.Tensor,
    template_dft: torch.Tensor,
    ctf_filters: torch.Tensor,
    whitening_filter_template: torch.Tensor,
    defocus_values: torch.Tensor,
    pixel_values: torch.Tensor,
    euler_angles: torch.Tensor,
    device: torch.device,
    checkpoint_path: str,
    orientation_batch_size: int = 1,
    checkpoint_interval_s: float = 600.0,
    max_batches: int | None = None,
    mag_matrix: torch.Tensor | None = None,
) -> dict[str, torch.Tensor] | None:
    """``core_match_template`` on one device with checkpoint / resume of the statistics.

    Implements the restart the reference lists as TODO
    (``pydantic_models/results/match_template_result.py:51-56``).  If
    ``checkpoint_path`` exists and belongs to the same search (same inputs and batch
    size) the search continues from its next orientation; the statistics are written
    every ``checkpoint_interval_s`` seconds and when the search stops.  With
    ``max_batches`` the call stops after that many orientation batches and returns
    None (the checkpoint then holds the progress); otherwise it returns the output
    dictionary of ``core_match_template``.  A resumed search gives the same maps, bit
    for bit, as an uninterrupted one.
    """
    devs = _as_device_list(device)
    if len(devs) != 1:
        raise ValueError("core_match_template_resumable runs on exactly one device")
    dev = devs[0]
    H = image_dft.shape[0]
    W = 2 * (image_dft.shape[1] - 1)
    h = template_dft.shape[1]
    O, D, C = euler_angles.shape[0], defocus_values.shape[0], pixel_values.shape[0]
    batch = max(1, min(orientation_batch_size, O))
    fp = search_fingerprint(image_dft, template_dft, ctf_filters, whitening_filter_template, defocus_values,
                            pixel_values, euler_angles, batch, mag_matrix)
    plan = MatchPlan(H, W, h, batch, D, C, dev)
    if mag_matrix is not None:
        plan.set_magnification(mag_matrix)
    with torch.cuda.device(dev):
        plan.set_inputs(image_dft, ctf_filters, whitening_filter_template)
        tdft = template_dft.to(dev, torch.complex64).contiguous()
        eul = euler_angles.to(dev, torch.float32).contiguous()
        plan._tdft, plan._eul = tdft, eul
        plan.reset_stats()
        start = 0
        if os.path.exists(checkpoint_path):
            start = plan.load_state(checkpoint_path, fp)
        last = time.monotonic()
        done = 0
        i = start
        while i < O:
            if max_batches is not None and done >= max_batches:
                break
            plan.step(tdft, eul, i, min(batch, O - i))
            i = min(O, i + batch)
            done += 1
            if time.monotonic() - last >= checkpoint_interval_s:
                plan.save_state(checkpoint_path, i, fp)
                last = time.monotonic()
        plan.save_state(checkpoint_path, i, fp)
        if i < O:
            return None
    return _finish(plan, defocus_values, O, D, C)


def raw_statistics(plan: MatchPlan) -> dict[str, torch.Tensor]:
    """Running maps of a plan as CPU tensors (sums as float64 hi+lo)."""
    s = [t.double().cpu() for t in plan.sums]
    return {
        "mip": plan.mip.cpu(),
        "best_global_index": plan.best_index.cpu(),
        "correlation_sum": s[0] + s[1],
        "correlation_squared_sum": s[2] + s[3],
    }
