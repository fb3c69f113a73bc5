"""Launch-shape autotuner for the search plan.

A ``MatchPlan`` carries launch-shape options (``tdtm_match_plan_set_option``): column-pass
columns per block, paired correlations, correlation chunks per column group, twiddle
placement and row-pass tile staging.  None of them changes the results (the kernels do
the same arithmetic in the same order), only the speed, and the best choice depends on
the image, template and batch shape.  ``autotune`` times a list of candidate option sets
on the plan's own inputs, keeps the fastest and caches the choice per
(device, H, W, h, B, D, C) in a JSON file.  The plan's running statistics are saved
before and restored after the timing steps, so tuning can run at any point of a search.

No counterpart in the reference (it has one fixed launch pattern per backend).
"""

from __future__ import annotations

import json
import os
import time

import torch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "leopard_em_amd", "autotune.json")

# option sets tried by default; the first is the library default
CANDIDATES: list[dict[str, int]] = [
    {},
    {"colpass_pair": 1},
    {"colpass_twiddles": 1},
    {"colpass_twiddles": 0},
    {"colpass_columns": 1},
    {"colpass_columns": 1, "colpass_pair": 0},
    {"rowpass_staging": 0},
    {"rowpass_twiddles": 0},
    {"rowpass_rows": 2},
    {"rowpass_rows": 4},
    {"colpass_store": 0},
    {"colpass_persistent": 0},
    {"colpass_store": 0, "colpass_persistent": 0},
]

_OPTION_DEFAULTS = {
    "colpass_columns": 2,
    "colpass_pair": 0,
    "colpass_chunks": 0,
    "colpass_twiddles": 2,
    "colpass_store": 1,
    "colpass_persistent": 1,
    "rowpass_twiddles": 1,
    "rowpass_staging": 3,
    "rowpass_rows": 1,
}


def _key(plan) -> str:
    name = torch.cuda.get_device_name(plan.device)
    return f"{name}|{plan.H}x{plan.W}|h{plan.h}|B{plan.B}|D{plan.D}|C{plan.C}"


def apply_options(plan, options: dict[str, int]) -> None:
    """Reset every tunable option to its default, then apply ``options`` (row-pass
    twiddles in LDS are switched off for the staging modes that leave no room)."""
    plan.set_option("rowpass_twiddles", 0)
    plan.set_option("rowpass_staging", _OPTION_DEFAULTS["rowpass_staging"])
    for k in ("colpass_columns", "colpass_pair", "colpass_chunks", "colpass_twiddles", "colpass_store",
              "colpass_persistent", "rowpass_rows"):
        plan.set_option(k, _OPTION_DEFAULTS[k])
    want = dict(options)
    tw = want.pop("rowpass_twiddles", None)
    staging = want.pop("rowpass_staging", _OPTION_DEFAULTS["rowpass_staging"])
    if staging in (1, 2):
        want["rowpass_rows"] = 4  # asynchronous staging copies whole tile rows
    for k, v in want.items():
        plan.set_option(k, v)
    plan.set_option("rowpass_staging", staging)
    if tw is None:
        tw = 0 if staging in (1, 2) else _OPTION_DEFAULTS["rowpass_twiddles"]
    plan.set_option("rowpass_twiddles", tw)


def _load_cache(path: str) -> dict:
    try:
        with open(path, encoding="utf-8") as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(plan, template_dft: torch.Tensor, euler_deg: torch.Tensor, steps: int = 3,
             candidates: list[dict[str, int]] | None = None, cache_path: str | None = DEFAULT_CACHE,
             force: bool = False) -> dict:
    """Pick and apply the fastest option set for ``plan``; returns
    ``{"options": ..., "ms_per_step": ..., "timings": [...], "cached": bool}``.

    The plan's inputs (``set_inputs``) must be set.  Timing uses ``steps`` full
    orientation batches from the start of ``euler_deg`` per candidate after one warm-up
    step.
    """
    key = _key(plan)
    if cache_path and not force:
        hit = _load_cache(cache_path).get(key)
        if hit is not None:
            apply_options(plan, hit["options"])
            return dict(hit, cached=True)
    cands = candidates if candidates is not None else CANDIDATES
    saved = [t.clone() for t in (plan.mip, plan.best_index, *plan.sums)]
    B = plan.B
    n_avail = euler_deg.shape[0]
    count = min(B, n_avail)
    timings = []
    stream = torch.cuda.current_stream(plan.device)
    try:
        for opts in cands:
            try:
                apply_options(plan, opts)
            except Exception as e:  # noqa: BLE001 -- an option set this size cannot use
                timings.append({"options": opts, "error": str(e)})
                continue
            plan.step(template_dft, euler_deg, 0, count)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for i in range(steps):
                start = (i * B) % max(1, n_avail - count + 1)
                plan.step(template_dft, euler_deg, start, count)
            e1.record(stream)
            e1.synchronize()
            timings.append({"options": opts, "ms_per_step": e0.elapsed_time(e1) / steps})
    finally:
        for dst, src in zip((plan.mip, plan.best_index, *plan.sums), saved):
            dst.copy_(src)
    ok = [t for t in timings if "ms_per_step" in t]
    if not ok:
        raise RuntimeError("autotune: no candidate option set ran")
    best = min(ok, key=lambda t: t["ms_per_step"])
    apply_options(plan, best["options"])
    result = {"options": best["options"], "ms_per_step": best["ms_per_step"], "timings": timings,
              "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    if cache_path:
        cache = _load_cache(cache_path)
        cache[key] = {k: result[k] for k in ("options", "ms_per_step", "tuned_at")}
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        tmp = f"{cache_path}.tmp{os.getpid()}"
        with open(tmp, "w", encoding="utf-8") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, cache_path)
    return dict(result, cached=False)
