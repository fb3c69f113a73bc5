"""2DTM search throughput on one MI355X: (orientation x defocus) correlations/s.

Workload (production shape): 4096 x 4096 synthetic micrograph, 512^3 Gaussian-blob
template (5000 blobs, sigma 1.5 px, radius 200 voxels, 1.0 A/px), 13 defocus planes
(-1200..+1200 A step 200), orientation batches of --batch orientations.  A "step" is
one orientation batch through the fused search (slice -> projection -> padded FFT
correlation -> MIP / argmax / sum / sum-sq update); inputs are resident in HBM.

Prints one JSON line (rank 0) with the throughput, an HBM roofline record for the
fused step and a CPU baseline (the float32 CPU oracle on a bounded sample of the same
workload, timed on this host).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "leopard-em_amd"))

HBM_PEAK_GBS = 8000.0
# Published H100 80 GB rates of the reference (docs/benchmarks.md:32, :42), by image (H, W)
REFERENCE_H100_CORR_PER_S = {(4096, 4096): 1650.8, (4092, 5760): 897.9}
TRAFFIC_SUMMARY = os.path.join(REPO, "profiles", "match_step_4096x4096_t512_b8_d13_traffic.md")


def algorithmic_bytes_per_correlation(H: int, W: int) -> int:
    """One complex64 half spectrum written and read back between the 2D inverse FFT passes."""
    return 2 * 8 * H * (W // 2 + 1)


def make_inputs(device, H, W, h, n_blobs, radius, n_particles, seed=0, return_micrograph=False):
    """Synthetic 2DTM search inputs built on ``device``: Gaussian-blob template,
    micrograph of embedded CTF-filtered projections at SNR 0.05 with per-particle defocus,
    whitened image spectrum, template rfft and the 13-plane CTF stack.  With
    ``return_micrograph`` the SyntheticMicrograph (true poses, centres, defocus) is
    appended to the returned tuple."""
    import torch

    from leopard_em_amd import synthetic
    from leopard_em_amd.utils import preprocessing as pp

    t = synthetic.make_blob_template(seed, h, n_blobs, radius)
    vol = t.volume(device=device)
    tdft = pp.volume_to_rfft_fourier_slice(vol).to(torch.complex64).contiguous()
    mg = synthetic.make_micrograph(t, seed, (H, W), n_particles, 12200.0, 11800.0, 30.0, snr=0.05,
                                   defocus_spread=3000.0, device=device)
    image = mg.image.to(device=device, dtype=torch.float32).contiguous()
    img_dft, wf = pp.prepare_image(image, h)
    dfs = torch.arange(-1200.0, 1201.0, 200.0)
    ctf = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, torch.tensor([0.0]), device=device)
    if return_micrograph:
        return img_dft, tdft, ctf, wf.to(device), dfs, vol, image, mg
    return img_dft, tdft, ctf, wf.to(device), dfs, vol, image


def time_preparation(image, vol, h, dfs, device, reps=5, cpu_baseline=False):
    """Device time of the once-per-search input preparation (image spectrum + whitening,
    template 3D rfft, CTF stack) and its rate against the minimum HBM traffic: the
    input read once and the output written once."""
    import torch

    from leopard_em_amd.utils import preprocessing as pp

    stream = torch.cuda.current_stream(device)

    def timed(fn):
        fn()
        torch.cuda.synchronize(device)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    H, W = image.shape
    n = vol.shape[0]
    out = {}
    jobs = {
        "prepare_image": (lambda: pp.prepare_image_device(image, h), 4 * H * W + 8 * H * (W // 2 + 1)),
        "volume_rfft": (lambda: pp.volume_rfft_device(vol), 4 * n ** 3 + 8 * n * n * (n // 2 + 1)),
        "ctf_stack": (lambda: pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, torch.tensor([0.0]),
                                                  device=device), 4 * dfs.numel() * h * (h // 2 + 1)),
    }
    for name, (fn, nbytes) in jobs.items():
        ms = timed(fn)
        gbs = nbytes / (ms / 1e3) / 1e9
        out[name] = {"ms": ms, "min_bytes": nbytes, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
    if cpu_baseline:
        # the CPU oracle's restatement of the reference's host preparation, all host threads
        sys.path.insert(0, REPO)
        from oracle import filters as of

        torch.set_num_threads(os.cpu_count() or 1)
        img_c, vol_c = image.cpu(), vol.cpu()
        t0 = time.perf_counter()
        of.prepare_image_dft(img_c, (h, h // 2 + 1))
        t1 = time.perf_counter()
        of.volume_to_rfft_fourier_slice(vol_c)
        t2 = time.perf_counter()
        out["prepare_image"]["cpu_ms"] = (t1 - t0) * 1e3
        out["volume_rfft"]["cpu_ms"] = (t2 - t1) * 1e3
        out["cpu_cores"] = os.cpu_count() or 1
    return out


def cpu_baseline(img_dft, tdft, ctf, wf, eul, n_orient, batch):
    """Float32 CPU oracle (reference algorithm) on n_orient orientations x all defocus."""
    import torch

    sys.path.insert(0, REPO)
    from oracle import geometry as og
    from oracle import search as osr

    cores = os.cpu_count() or 1
    torch.set_num_threads(cores)
    X = img_dft.cpu()
    T = tdft.cpu()
    F = (ctf.cpu() * wf.cpu()[None, None])
    e = eul[:n_orient].cpu()
    # one warm-up correlation batch, then the timed sample
    osr.batched_cross_correlate(X, T, og.euler_to_rotmat_zyz(e[:1], torch.float32), F[:, :1])
    t0 = time.perf_counter()
    n = 0
    for i in range(0, n_orient, batch):
        rot = og.euler_to_rotmat_zyz(e[i:i + batch], torch.float32)
        cc = osr.batched_cross_correlate(X, T, rot, F)
        n += cc.shape[0] * cc.shape[1] * cc.shape[2]
        del cc
    dt = time.perf_counter() - t0
    return n / dt, cores, n, dt


def _refine_stack_inputs(dev, n, box, h, seed=0):
    """Refinement stack: ``n`` box x box crops cut around embedded particles
    of 8 synthetic micrographs (seeded sizes 2048..4092 x 2048..5760, CTF defocus spread
    +/-3000 A, SNR 0.05), each crop with its own local grid around its true pose perturbed
    by U(-2, 2) deg: orientation offsets +/-hw at 0.25 deg with hw in {1.5, 1.75, 2.0}
    (13^3, 15^3 or 17^3 offsets), defocus offsets +/-300 A at 50 A (13) and 1..9 pixel-size
    offsets 0.005 A apart."""
    import numpy as np
    import torch

    from leopard_em_amd import synthetic

    rng = np.random.default_rng(seed)
    t = synthetic.make_blob_template(seed, h, max(100, int(5000 * (h / 512) ** 3)), 200.0 * h / 512)
    vol = t.volume(device=dev)
    from leopard_em_amd.utils import preprocessing as pp

    tdft = pp.volume_to_rfft_fourier_slice(vol).to(torch.complex64).contiguous()
    per = [n // 8 + (1 if i < n % 8 else 0) for i in range(8)]
    crops, poses, truth, du, dv = [], [], [], [], []
    for m in range(8):
        if per[m] == 0:
            continue
        H = int(rng.integers(2048, 4093)) // 2 * 2
        W = int(rng.integers(2048, 5761)) // 2 * 2
        mg = synthetic.make_micrograph(t, seed + 31 * m, (H, W), per[m], 12200.0, 11800.0, 30.0, snr=0.05,
                                       defocus_spread=3000.0, margin=box // 2 + 1, device=dev)
        img = mg.image.to(dev)
        for i in range(per[m]):
            cy, cx = (int(round(v)) for v in mg.centers_yx[i])
            crop = img[cy - box // 2:cy + box // 2, cx - box // 2:cx + box // 2]
            crop = (crop - crop.mean()) / crop.std()
            X = torch.fft.rfft2(crop)
            X[0, 0] = 0
            crops.append(X.to(torch.complex64))
            truth.append(mg.eulers[i])
            poses.append(mg.eulers[i] + rng.uniform(-2.0, 2.0, 3))
            du.append(float(mg.defocus[i]) + 200.0)
            dv.append(float(mg.defocus[i]) - 200.0)
    stack = torch.stack(crops)
    offs, dfos, pxos = [], [], []
    for _ in range(n):
        hw = float(rng.choice([1.5, 1.75, 2.0]))
        a = torch.arange(-hw, hw + 1e-4, 0.25)
        offs.append(torch.stack(torch.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3))
        dfos.append(torch.arange(-300.0, 300.1, 50.0))
        nc = int(rng.integers(1, 10))
        pxos.append((torch.arange(nc, dtype=torch.float32) - (nc - 1) / 2) * 0.005)
    return (stack, tdft, torch.tensor(np.array(poses), dtype=torch.float32), offs, dfos,
            torch.tensor(du), torch.tensor(dv), torch.full((n,), 30.0), pxos, np.array(truth))


def bench_refine(args) -> dict:
    """Batched refinement throughput on crops of synthetic micrographs (``_refine_stack_inputs``):
    2000 crops of 512^2 against a 256^3 template (valid region 257^2), each crop with its
    own grid, batches of 32 offsets, one ``refine_template_ragged`` call (tdtm_refine_batch,
    full batches replayed from HIP graphs).  ``--refine-particles`` crops are timed between
    HIP events (default a 24-crop sample, extrapolated to the 2000-crop job); mean / std
    maps are 0 / 1 (z = cc), inputs resident on the device."""
    import numpy as np
    import torch

    from leopard_em_amd.backend import refine_template_ragged
    sys.path.insert(0, REPO)

    dev = torch.device("cuda:0")
    n, box, h = args.refine_particles, 512, 256
    print(f"[bench] refinement stack: {n} crops of {box}^2 from 8 micrographs, template {h}^3",
          file=sys.stderr, flush=True)
    stack, tdft, poses, offs, dfos, du, dv, da, pxos, truth = _refine_stack_inputs(dev, n, box, h)
    crop = (box - h + 1, box - h + 1)
    mean = torch.zeros((n,) + crop, device=dev)
    std = torch.ones((n,) + crop, device=dev)
    filt = torch.ones((n, h, h // 2 + 1), device=dev)
    kw = dict(voltage=300.0, spherical_aberration=2.7, amplitude_contrast_ratio=0.07, phase_shift=0.0,
              pixel_size=1.0, ctf_B_factor=0.0)

    def run(sl):
        return refine_template_ragged(stack[sl], tdft, poses[sl], offs[sl], dfos[sl], du[sl], dv[sl], da[sl],
                                      pxos[sl], mean[sl], std[sl], kw, filt[sl], dev,
                                      batch_size=args.refine_batch)

    run(slice(0, 1))  # warm-up (plan, graph capture)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    res = run(slice(0, n))
    torch.cuda.synchronize(dev)
    sec = time.perf_counter() - t0
    counts = [int(o.shape[0]) * int(d.numel()) * int(p.numel()) for o, d, p in zip(offs, dfos, pxos)]
    ncorr = sum(counts)
    rate = ncorr / sec
    # angular error of the refined pose against the true pose (rotation angle of R_ref^T R_true)
    from leopard_em_amd.synthetic import _rotmat_zyz_np

    errs = []
    for i in range(n):
        r1 = _rotmat_zyz_np(*[float(v) for v in res["refined_euler_angles"][i]])
        r0 = _rotmat_zyz_np(*[float(v) for v in truth[i]])
        c = (np.trace(r1.T @ r0) - 1) / 2
        errs.append(float(np.degrees(np.arccos(np.clip(c, -1, 1)))))
    # the 2000-crop job: average correlations per crop of this generator x 2000
    mean_corr = float(np.mean([((2 * hw / 0.25 + 1) ** 3) for hw in (1.5, 1.75, 2.0)]) * 13 * 5)
    print(f"[bench] refine {n} crops: {sec:.2f} s, {rate:.0f} corr/s, median angular error "
          f"{np.median(errs):.2f} deg", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        # the float32 CPU oracle (reference algorithm) on one crop with a 5^3 offset grid x 13 defocus
        from oracle import search as osr

        cores = os.cpu_count() or 1
        torch.set_num_threads(cores)
        g3 = torch.tensor([-0.5, -0.25, 0.0, 0.25, 0.5])
        offs3 = torch.stack(torch.meshgrid(g3, g3, g3, indexing="ij"), -1).reshape(-1, 3)
        dfo3 = torch.arange(-300.0, 300.1, 50.0)
        pxo3 = torch.tensor([0.0])
        cargs = (stack[:1].cpu(), tdft.cpu(), poses[:1], offs3, dfo3, du[:1], dv[:1], da[:1], pxo3, mean[:1].cpu(),
                 std[:1].cpu(), kw, filt[:1].cpu())
        t0 = time.perf_counter()
        osr.core_refine_template(*cargs, batch_size=32)
        dt = time.perf_counter() - t0
        nc = offs3.shape[0] * dfo3.numel() * pxo3.numel()
        cpu = {"value": nc / dt, "unit": "correlations/s", "cores": cores, "kind": "port",
               "sample": f"1 crop x {offs3.shape[0]} offsets x {dfo3.numel()} defocus = {nc} correlations in "
                         f"{dt:.1f} s, torch CPU float32 oracle"}
        print(f"[bench] CPU baseline: {nc / dt:.1f} corr/s", file=sys.stderr, flush=True)
    return {
        "metric": "refinement correlations/s (512^2 crops x 256^3 template, per-crop grids)",
        "value": rate,
        "unit": "correlations/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (crops of seeded synthetic micrographs with embedded particles; mean 0 / std 1 maps)",
        "config": {"workload": f"batched ragged refinement, {n} crops of {box}x{box} from 8 micrographs "
                               f"(2048..4092 x 2048..5760), {h}^3 template, per-crop grids: 13^3 / 15^3 / 17^3 "
                               "offsets at 0.25 deg x 13 defocus x 1..9 pixel sizes, batches of 32 offsets",
                   "crops_timed": n, "correlations_timed": ncorr, "seconds": sec,
                   "crops_per_s": n / sec,
                   "job_2000_crops_s_extrapolated": 2000 * mean_corr / rate,
                   "median_angular_error_deg": float(np.median(errs))},
        "cpu_baseline": cpu,
    }


def bench_proposal(args) -> dict:
    """Learned orientation proposals (leopard_em_amd.proposal.run_experiment): the network
    trained on the GPU by its own kernels on the plain search of ``--proposal-train``
    seeded 512^2 micrographs (64^3 template, 10 / 5 deg grid, ``--proposal-bin-deg`` bins,
    ``--proposal-tile`` tiles), the guided search on ``--proposal-test`` held-out
    micrographs against the exhaustive search."""
    import torch

    from leopard_em_amd import proposal

    dev = torch.device("cuda:0")
    out = proposal.run_experiment(dev, n_train=args.proposal_train, n_test=args.proposal_test,
                                  n_particles=args.proposal_particles, epochs=40, tile=args.proposal_tile,
                                  bin_deg=args.proposal_bin_deg,
                                  operating_points=[(0.3, 0.9), (2.0, 0.9), (0.0, 2.0)],
                                  log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    ex_rate = out["orientations"] * out["test_micrographs"] / out["exhaustive_s"]
    return {
        "metric": "proposal-guided search: (orientation x defocus)/s equivalent and recall of z > 7.5 peaks",
        "value": out["orientations"] * out["test_micrographs"] / out["guided_s"],
        "unit": "orientations/s (exhaustive-equivalent, 512^2 micrographs)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 network (f32 accumulation), f32 search",
        "data": f"synthetic (seeded micrographs, {args.proposal_particles} embedded particles each, SNR 0.3)",
        "config": {"workload": f"learned orientation proposal, {out['train_micrographs']} training / "
                               f"{out['test_micrographs']} held-out micrographs 512^2, 64^3 template, "
                               f"{out['orientations']:,} orientations, {out['bins']} bins at {out['bin_deg']:g} deg, "
                               f"top 5% bins per {out['tile']}^2 tile",
                   "exhaustive_orientations_per_s": ex_rate},
        "proposal": out,
        "cpu_baseline": None,
    }


SEARCH_SOURCES = ("common.hpp", "fft.hpp", "fft4.hpp", "fft_mfma.hpp", "match.hip")


def kernel_sources_digest() -> str:
    """SHA-256 over the HIP sources of the search step (match.hip and the headers it
    includes; names and contents in this order)."""
    import hashlib

    csrc = os.path.join(REPO, "leopard-em_amd", "csrc")
    h = hashlib.sha256()
    for name in SEARCH_SOURCES:
        h.update(name.encode())
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def read_summary_field(path: str, key: str):
    try:
        with open(path, encoding="utf-8") as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        return None
    return None


def measure_recall(plan, eul, dfs, total, mg, h, n_true, zcut, radius=2.0):
    """Embedded particles found by the search just run: z = (MIP - mean) / std over the
    timed orientations, peaks above ``zcut`` on the valid crop picked with the reference's
    greedy 5-pixel suppression (analysis/zscore_metric.py:42-83); a particle whose true
    pose is among the orientations counts as found when a peak lies within ``radius``
    pixels of its template top-left (centre - h//2)."""
    import numpy as np
    import torch

    from leopard_em_amd.analysis import find_peaks_from_zscore

    dvals = dfs.to(plan.device, torch.float32).contiguous()
    maps = plan.finalize(eul, dvals, total)
    vy, vx = plan.H - h + 1, plan.W - h + 1
    z = maps["scaled_mip"][:vy, :vx].cpu()
    py, px = find_peaks_from_zscore(z, zcut)
    peaks = np.stack([py.numpy(), px.numpy()], 1).astype(float) if len(py) else np.zeros((0, 2))
    tl = np.rint(mg.centers_yx[:n_true]).astype(int) - h // 2
    inside = (tl[:, 0] >= 0) & (tl[:, 0] < vy) & (tl[:, 1] >= 0) & (tl[:, 1] < vx)
    tl = tl[inside]
    found = 0
    matched = np.zeros(len(peaks), dtype=bool)
    for t in tl:
        if len(peaks):
            d = np.hypot(peaks[:, 0] - t[0], peaks[:, 1] - t[1])
            hit = d <= radius
            if hit.any():
                found += 1
                matched |= hit
    zt = [float(z[t[0], t[1]]) for t in tl]
    return {"zcut": zcut, "particles": int(len(tl)), "found": int(found),
            "recall": found / len(tl) if len(tl) else None, "false_peaks": int((~matched).sum()),
            "median_particle_z": float(np.median(zt)) if zt else None,
            "rule": f"z > {zcut} peak within {radius} px of the true template top-left; the true poses are in the "
                    "timed orientation slice"}


def read_traffic(path: str, key: str = "hbm_bytes_per_step"):
    """HBM bytes (``key``) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    try:
        with open(path, encoding="utf-8") as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return int(ln.split(":", 1)[1].strip())
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 512 steps of 8 orientations: the 4096-orientation slice of the production grid
    ap.add_argument("--steps", type=int, default=512)
    # untimed steps first: the clocks settle within a few steps
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8, help="orientations per step")
    ap.add_argument("--size", type=int, default=4096, help="image edge (H = W)")
    ap.add_argument("--image", default=None,
                    help="image HxW, e.g. 4092x5760 (K3; run in a 4096 x 8192 grid); overrides --size")
    ap.add_argument("--template", type=int, default=512)
    ap.add_argument("--cpu-orientations", type=int, default=16, help="CPU baseline sample (x 13 defocus)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--inputs", choices=("synthetic", "random"), default="synthetic",
                    help="random: skip micrograph synthesis (timing-equivalent; for profiler runs)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (reference kernel for HBM byte counters)")
    ap.add_argument("--no-phase-timing", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="time plain step launches instead of graph replays")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="launch-shape option of the plan (tdtm_match_plan_set_option), repeatable")
    ap.add_argument("--fft-engine", type=int, default=0,
                    help="column transforms: 0 FP32 VALU; matrix cores 1 fp16 split, 2 bf16 split, "
                         "3 fp16, 4 bf16 (4096-row grids)")
    ap.add_argument("--zcut", type=float, default=7.5, help="z-score cut-off of the recall measurement")
    ap.add_argument("--no-variants", action="store_true", help="skip timing the matrix-core column-pass variants")
    ap.add_argument("--no-prep-timing", action="store_true",
                    help="skip timing the input preparation kernels (synthetic inputs only)")
    ap.add_argument("--workload", choices=("match", "refine", "proposal"), default="match",
                    help="refine: particle refinement throughput; proposal: learned orientation proposals")
    ap.add_argument("--refine-particles", type=int, default=24,
                    help="crops timed by --workload refine (2000 = the whole 8-micrograph job)")
    ap.add_argument("--refine-batch", type=int, default=32, help="orientation offsets per refinement batch")
    ap.add_argument("--proposal-train", type=int, default=64, help="training micrographs of --workload proposal")
    ap.add_argument("--proposal-test", type=int, default=16, help="held-out micrographs of --workload proposal")
    ap.add_argument("--proposal-particles", type=int, default=6, help="particles per proposal micrograph")
    ap.add_argument("--proposal-tile", type=int, default=256, help="tile edge (multiple of 64) of the proposals")
    ap.add_argument("--proposal-bin-deg", type=float, default=30.0, help="coarse orientation bin step (deg)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU sharding is a design note, "
              "not a measured configuration", file=sys.stderr)
        sys.exit(2)
    if args.workload == "refine":
        print(json.dumps(bench_refine(args)), flush=True)
        return
    if args.workload == "proposal":
        print(json.dumps(bench_proposal(args)), flush=True)
        return

    import torch

    from leopard_em_amd.backend import MatchPlan
    from leopard_em_amd.utils.search import get_uniform_euler_angles

    dev = torch.device("cuda:0")
    H = W = args.size
    if args.image:
        H, W = (int(v) for v in args.image.lower().split("x"))
    h = args.template
    print(f"[bench] building {args.inputs} inputs {H}x{W}, template {h}^3", file=sys.stderr, flush=True)
    if args.inputs == "synthetic":
        n_blobs = 5000 if h == 512 else max(100, int(5000 * (h / 512) ** 3))
        img_dft, tdft, ctf, wf, dfs, vol, image, mg = make_inputs(dev, H, W, h, n_blobs, 200.0 * h / 512, 200,
                                                                  return_micrograph=True)
    else:
        mg = None
        gen = torch.Generator(device=dev).manual_seed(0)
        img_dft = torch.randn((H, W // 2 + 1), dtype=torch.complex64, device=dev, generator=gen)
        tdft = torch.randn((h, h, h // 2 + 1), dtype=torch.complex64, device=dev, generator=gen)
        dfs = torch.arange(-1200.0, 1201.0, 200.0)
        ctf = torch.rand((1, dfs.shape[0], h, h // 2 + 1), device=dev, generator=gen)
        wf = torch.rand((h, h // 2 + 1), device=dev, generator=gen)
    D = dfs.shape[0]
    B = args.batch
    n_timed = args.steps * B
    # orientations: the embedded particles' true poses first (synthetic inputs), then a
    # strided slice of the production grid (out-of-plane 2.5 deg, in-plane 1.5 deg); the
    # warm-up steps take other grid orientations
    grid = get_uniform_euler_angles(psi_step=1.5, theta_step=2.5)
    full_grid = grid.shape[0]
    n_grid = n_timed + args.warmup * B
    stride = max(1, full_grid // max(n_grid, 1))
    strided = grid[::stride][:n_grid]
    true_poses = torch.tensor(mg.eulers, dtype=torch.float32) if mg is not None else torch.zeros((0, 3))
    n_true = min(true_poses.shape[0], n_timed)
    eul = torch.cat([true_poses[:n_true], strided[:n_timed - n_true]]).contiguous().to(dev)
    eul_warm = strided[n_timed - n_true:][:max(args.warmup * B, 1)].contiguous().to(dev)

    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize(dev)
        b.copy_(a)
        torch.cuda.synchronize(dev)
        print(f"[bench] copy calibration: {nbytes} bytes read + written", file=sys.stderr, flush=True)
        del a, b

    plan = MatchPlan(H, W, h, B, D, 1, dev)
    plan.set_option("graph", 0 if args.no_graph else 1)
    if args.fft_engine:
        plan.set_option("fft_engine", args.fft_engine)
    for opt in args.option:
        key, val = opt.split("=", 1)
        plan.set_option(key, int(val))
    plan.set_inputs(img_dft, ctf, wf)
    plan.reset_stats()
    stream = torch.cuda.current_stream(dev)
    plan.run(tdft, eul_warm, 0, min(args.warmup * B, eul_warm.shape[0]))
    torch.cuda.synchronize(dev)
    plan.reset_stats()
    torch.cuda.synchronize(dev)
    print("[bench] warm-up done", file=sys.stderr, flush=True)

    # timed region: the whole orientation slice through tdtm_match_run (full batches
    # replayed from the captured step graph unless --no-graph)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    ev0.record(stream)
    plan.run(tdft, eul, 0, n_timed)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    corr = n_timed * D
    value = corr / (gpu_ms / 1e3)
    ab = algorithmic_bytes_per_correlation(H, W)
    achieved = ab * B * D / (ms_per_step / 1e3) / 1e9
    print(f"[bench] {args.steps} steps ({n_timed} orientations): {gpu_ms:.1f} ms (wall {wall * 1e3:.1f} ms), "
          f"{value:.1f} corr/s", file=sys.stderr, flush=True)

    recall = None
    if mg is not None:
        recall = measure_recall(plan, eul, dfs, n_timed * D, mg, h, n_true, args.zcut)
        print(f"[bench] recall at z > {args.zcut}: {recall['found']} of {recall['particles']} embedded particles, "
              f"{recall['false_peaks']} other peaks", file=sys.stderr, flush=True)

    # the same batches launched as plain steps (no graph), for the per-step launch cost
    n_cmp = min(args.steps, 64)
    plan.set_option("graph", 0)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    plan.run(tdft, eul, 0, n_cmp * B)
    e1.record(stream)
    e1.synchronize()
    ms_direct = e0.elapsed_time(e1) / n_cmp
    plan.set_option("graph", 1)
    e0.record(stream)
    plan.run(tdft, eul, 0, n_cmp * B)
    e1.record(stream)
    e1.synchronize()
    ms_graph = e0.elapsed_time(e1) / n_cmp
    print(f"[bench] ms/step over {n_cmp} steps: direct launches {ms_direct:.3f}, graph replays {ms_graph:.3f}",
          file=sys.stderr, flush=True)

    # the matrix-core column-pass variants on the same plan and orientations (4096-row grids):
    # throughput of plain steps and the column pass's phase time
    variants = None
    if plan.Hg == 4096 and args.fft_engine == 0 and not args.no_variants:
        variants = {}
        names = {1: "MFMA fp16, split hi/lo data and matrices", 2: "MFMA bf16, split hi/lo data and matrices"}
        for e, name in names.items():
            plan.set_option("fft_engine", e)
            plan.run(tdft, eul, 0, 2 * B)
            e0.record(stream)
            plan.run(tdft, eul, 0, n_cmp * B)
            e1.record(stream)
            e1.synchronize()
            ms_v = e0.elapsed_time(e1) / n_cmp
            plan.set_option("phase_timing", 1)
            for i in range(min(n_cmp, 16)):
                plan.step(tdft, eul, i * B, B)
            ph, nph = plan.phase_times()
            plan.set_option("phase_timing", 0)
            variants[f"fft_engine={e}"] = {"name": name, "ms_per_step": ms_v, "value": B * D / (ms_v / 1e3),
                                           "colpass_ms": ph["colpass"] / nph, "steps": n_cmp}
            print(f"[bench] fft_engine {e} ({name}): {ms_v:.3f} ms/step, colpass {ph['colpass'] / nph:.3f} ms",
                  file=sys.stderr, flush=True)
        plan.set_option("fft_engine", 0)

    phases = None
    if not args.no_phase_timing:
        # HIP events between the phases of each step (plain launches)
        plan.set_option("phase_timing", 1)
        for i in range(n_cmp):
            plan.step(tdft, eul, i * B, B)
        ph, n = plan.phase_times()
        plan.set_option("phase_timing", 0)
        phases = {k: v / n for k, v in ph.items()}
        print("[bench] phase ms/step: " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()),
              file=sys.stderr, flush=True)

    measured_shape = (H, W, h, B, D) == (4096, 4096, 512, 8, 13)
    # the committed counter summary counts only if it was measured on these kernel sources
    traffic_current = read_summary_field(TRAFFIC_SUMMARY, "kernel_sources_sha256") == kernel_sources_digest()
    traffic = read_traffic(TRAFFIC_SUMMARY) if measured_shape and traffic_current else None
    # the dominant kernel on its own: k_colpass reads the row spectra (h rows per
    # correlation) and the image spectrum, and writes the half-spectrum intermediate
    # (all grid rows), W/2+1 complex64 columns each; its time from the same phase events
    dominant = None
    if phases:
        wr = plan.Wg // 2 + 1
        cp_bytes = (B * D * h * wr + plan.Hg * wr + B * D * plan.Hg * wr) * 8
        cp_gbs = cp_bytes / (phases["colpass"] / 1e3) / 1e9
        dominant = {"kernel": "k_colpass", "ms": phases["colpass"], "algorithmic_bytes": cp_bytes,
                    "achieved": cp_gbs, "frac": cp_gbs / HBM_PEAK_GBS,
                    "traffic": (read_traffic(TRAFFIC_SUMMARY, "colpass_bytes_per_launch")
                                if measured_shape and traffic_current else None)}
    ref_rate = REFERENCE_H100_CORR_PER_S.get((H, W))

    prep = None
    if args.inputs == "synthetic" and not args.no_prep_timing:
        prep = time_preparation(image, vol, h, dfs, dev, cpu_baseline=not args.no_cpu_baseline)
        print("[bench] preparation ms: " + ", ".join(f"{k} {v['ms']:.3f}" for k, v in prep.items()
                                                     if isinstance(v, dict)),
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (float32 oracle) ...", file=sys.stderr, flush=True)
        rate, cores, n, dt = cpu_baseline(img_dft, tdft, ctf, wf, eul, args.cpu_orientations, args.cpu_orientations)
        cpu = {"value": rate, "unit": "correlations/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_orientations} orientations x {D} defocus = {n} correlations of "
                         f"{H}x{W} x {h}^3 in {dt:.1f} s, torch CPU float32 oracle"}

    line = {
        "metric": ("(orientation x defocus) correlations/s on 4096^2 image" if (H, W) == (4096, 4096)
                   else f"(orientation x defocus) correlations/s on {H}x{W} image"),
        "value": value,
        "unit": "correlations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_direct_launch": ms_direct,
        "ms_per_step_graph_replay": ms_graph,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / ref_rate if ref_rate else None,
        "baseline": ({"value": ref_rate, "hardware": "H100 80 GB", "source": "reference docs/benchmarks.md:32-42"}
                     if ref_rate else None),
        "dtype": "f32",
        "data": ("synthetic (seeded Gaussian-blob template and micrograph; random orientations)"
                 if args.inputs == "synthetic" else "synthetic (random spectra; random orientations)"),
        "config": {
            "workload": f"plain 2DTM search, {H}x{W} micrograph x {h}^3 template, {D} defocus, "
                        f"{B} orientations per step",
            "image": [H, W], "search_grid": [plan.Hg, plan.Wg], "template": h, "defocus": D,
            "orientation_batch": B,
            "orientations": (f"{n_timed} = {n_true} embedded particles' true poses + a strided slice of the "
                             "2.5/1.5 deg uniform grid"),
            "full_grid_orientations": full_grid,
            "full_grid_hours_extrapolated": full_grid * D / value / 3600.0,
            **({"options": dict(o.split("=", 1) for o in args.option)} if args.option else {}),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "tdtm_match_step launch sequence (k_slice, k_proj_*, k_rowfwd, k_colpass, k_rowpass)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"{os.path.relpath(TRAFFIC_SUMMARY, REPO)} (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                               "measured on these kernel sources)" if traffic is not None else
                               "none: the committed counter summary is for other kernel sources or shapes"),
            "algorithmic_bytes_per_step": ab * B * D,
            "phases_ms_per_step": phases,
            "dominant_kernel": dominant,
        },
        "recall": recall,
        "variants": variants,
        "cpu_baseline": cpu,
        "preparation": prep,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
