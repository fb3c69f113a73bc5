"""Learned orientation proposals for the 2DTM search (no counterpart in the reference).

A small CNN (``tdtm_proposal_*`` in ``csrc/proposal.hip``: four 3x3 conv layers on 64 x 64
inputs, an MLP head) scores coarse orientation bins for each tile of a micrograph (256 x
256 by default, averaged to 64 x 64), plus one "empty tile" class.  It is trained on the best-orientation maps the
plain search produces on seeded synthetic micrographs.  The proposal-guided search then
visits, per tile, only the fine orientations inside the tile's top-k bins (plus a fixed
random subset used for the per-pixel mean / std), falls back to the exhaustive search for
tiles whose top-1 probability is below a threshold, and skips tiles the network calls
empty with high confidence.  Every reported peak comes from an exact correlation of the
guided search; recall is measured against the exhaustive search's z-score peaks.

Pieces:
* ``ProposalNet``: the HIP network (forward, backward, Adam) through the C ABI.
* ``OrientationBins``: fine orientations -> coarse bins of the uniform grid generator.
* ``tile_grid`` / ``tile_inputs``: tiling of a micrograph and the 64 x 64 network inputs.
* ``training_examples``: tiles and targets from the plain search's peaks.
* ``guided_search``: the proposal-guided search of one micrograph.
* ``peak_recall``: recall of the exhaustive z-score peaks.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .utils.search import get_uniform_euler_angles

TILE = 256        # default micrograph tile edge seen by the network
NET_IN = 64       # network input edge: a tile is averaged TILE / NET_IN x TILE / NET_IN
CHANNELS = (1, 16, 32, 64, 64)
HIDDEN = 256
FEATURES = 4 * 4 * 64


def param_shapes(num_classes: int):
    out = []
    for i in range(4):
        out += [(CHANNELS[i + 1], 3, 3, CHANNELS[i]), (CHANNELS[i + 1],)]
    return out + [(HIDDEN, FEATURES), (HIDDEN,), (num_classes, HIDDEN), (num_classes,)]


def init_params(num_classes: int, seed: int = 0) -> torch.Tensor:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases, in the ABI's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in param_shapes(num_classes):
        if len(s) == 1:
            parts.append(np.zeros(s, np.float32))
        else:
            fan_in = int(np.prod(s[1:]))
            parts.append((rng.standard_normal(s) * math.sqrt(2.0 / fan_in)).astype(np.float32))
    return torch.from_numpy(np.concatenate([p.reshape(-1) for p in parts]))


class ProposalNet:
    """The HIP proposal network on one device."""

    def __init__(self, num_classes: int, max_batch: int, device: torch.device, params: torch.Tensor | None = None,
                 seed: int = 0):
        if device.type != "cuda":
            raise _lib.TdtmError(f"the proposal network runs on a HIP device, got {device}")
        self.lib = _lib.load()
        self.device, self.L, self.max_batch = device, num_classes, max_batch
        h = _lib.c_void_p()
        with torch.cuda.device(device):
            _lib.check(self.lib.tdtm_proposal_create(num_classes, max_batch, _lib.ctypes.byref(h)))
        self.handle = h
        self.nparam = int(self.lib.tdtm_proposal_num_params(h))
        self.set_params(init_params(num_classes, seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.tdtm_proposal_destroy(h)
            except Exception:  # noqa: BLE001 -- interpreter shutdown
                pass
            self.handle = None

    def _st(self):
        return _lib.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_params(self, params: torch.Tensor) -> None:
        p = params.to(self.device, torch.float32).contiguous()
        if p.numel() != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.numel()}")
        _lib.check(self.lib.tdtm_proposal_set_params(self.handle, _lib.ptr(p), self._st()))
        torch.cuda.current_stream(self.device).synchronize()

    def params(self) -> torch.Tensor:
        p = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.tdtm_proposal_get_params(self.handle, _lib.ptr(p), None, self._st()))
        return p

    def grads(self) -> torch.Tensor:
        g = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.tdtm_proposal_get_params(self.handle, None, _lib.ptr(g), self._st()))
        return g

    def forward(self, tiles: torch.Tensor) -> torch.Tensor:
        """Logits (N, L) of (N, 64, 64) input tiles, in batches of at most max_batch."""
        x = tiles.to(self.device, torch.float32).contiguous()
        out = torch.empty((x.shape[0], self.L), dtype=torch.float32, device=self.device)
        for i in range(0, x.shape[0], self.max_batch):
            n = min(self.max_batch, x.shape[0] - i)
            _lib.check(self.lib.tdtm_proposal_forward(self.handle, _lib.ptr(x[i:i + n]), n, _lib.ptr(out[i:i + n]),
                                                      self._st()))
        return out

    def backward(self, targets: torch.Tensor) -> torch.Tensor:
        """Per-sample loss of the last forward batch; gradients kept in the network."""
        t = targets.to(self.device, torch.float32).contiguous()
        loss = torch.empty(t.shape[0], dtype=torch.float32, device=self.device)
        _lib.check(self.lib.tdtm_proposal_backward(self.handle, _lib.ptr(t), t.shape[0], _lib.ptr(loss), self._st()))
        return loss

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        _lib.check(self.lib.tdtm_proposal_adam_step(self.handle, lr, beta1, beta2, eps, self._st()))

    def train(self, tiles: torch.Tensor, targets: torch.Tensor, epochs: int, lr: float = 1e-3, batch: int = 64,
              seed: int = 0) -> list[float]:
        """Adam on shuffled mini-batches; returns the mean loss of every epoch."""
        g = torch.Generator().manual_seed(seed)
        x = tiles.to(self.device, torch.float32)
        t = targets.to(self.device, torch.float32)
        hist = []
        for _ in range(epochs):
            perm = torch.randperm(x.shape[0], generator=g).to(self.device)
            tot = 0.0
            for i in range(0, x.shape[0], batch):
                idx = perm[i:i + batch]
                self.forward(x[idx])
                tot += float(self.backward(t[idx]).sum())
                self.adam_step(lr)
            hist.append(tot / x.shape[0])
        return hist


# ----------------------------------------------------------------------- orientation bins
class OrientationBins:
    """Coarse bins of the uniform orientation grid: bin b is row b of
    ``get_uniform_euler_angles(bin_deg, bin_deg)`` (theta rings, phi values per ring, psi
    values), and a fine orientation (phi, theta, psi) falls into the ring nearest its
    theta, the ring's phi nearest its phi and the psi nearest its psi.  The last class of
    the network (index ``num_bins``) is "empty tile"."""

    def __init__(self, bin_deg: float = 30.0):
        self.bin_deg = bin_deg
        coarse = get_uniform_euler_angles(psi_step=bin_deg, theta_step=bin_deg).double().numpy()
        self.thetas = np.unique(np.round(coarse[:, 1], 6))
        self.n_psi = int(round(360.0 / bin_deg))
        self.n_phi = np.array([int(np.sum(np.isclose(coarse[:, 1], t))) // self.n_psi for t in self.thetas])
        self.ring_start = np.concatenate([[0], np.cumsum(self.n_phi * self.n_psi)[:-1]])
        self.num_bins = int(coarse.shape[0])
        self.num_classes = self.num_bins + 1
        self.empty = self.num_bins

    def of(self, euler_deg: torch.Tensor) -> np.ndarray:
        """Bin of every (phi, theta, psi) row."""
        e = euler_deg.double().cpu().numpy()
        ring = np.clip(np.rint(e[:, 1] / self.bin_deg).astype(int), 0, len(self.thetas) - 1)
        nphi = self.n_phi[ring]
        iphi = np.rint((e[:, 0] % 360.0) / (360.0 / nphi)).astype(int) % nphi
        ipsi = np.rint((e[:, 2] % 360.0) / self.bin_deg).astype(int) % self.n_psi
        return self.ring_start[ring] + iphi * self.n_psi + ipsi


# ----------------------------------------------------------------------------- tiling
def tile_grid(H: int, W: int, h: int, tile: int = TILE):
    """Tiles of a micrograph: [(y0, x0, (oy0, oy1), (ox0, ox1))] where the tile's crop is
    image[y0 : y0 + tile, x0 : x0 + tile] and it owns the template positions (top-left
    corners) oy0 <= y < oy1, ox0 <= x < ox1 of the valid region (H - h + 1) x (W - h + 1).
    Tile starts step by tile / 2; the last tile of a row / column sits at the image edge
    and owns the rest."""
    def starts(n):
        vmax = n - h + 1  # valid positions
        s = list(range(0, max(1, n - tile + 1), tile // 2))
        if s[-1] != n - tile:
            s.append(n - tile)
        own = [(a, s[i + 1] if i + 1 < len(s) else vmax) for i, a in enumerate(s)]
        return s, own

    ys, oys = starts(H)
    xs, oxs = starts(W)
    return [(y, x, oy, ox) for y, oy in zip(ys, oys) for x, ox in zip(xs, oxs)]


def tile_inputs(image: torch.Tensor, tiles, tile: int = TILE) -> torch.Tensor:
    """(n_tiles, 64, 64) network inputs: each tile x tile crop averaged over
    (tile / 64)^2 pixels, zero mean, unit std."""
    if tile % NET_IN:
        raise ValueError(f"tile edge {tile} is not a multiple of {NET_IN}")
    out = []
    for y0, x0, _, _ in tiles:
        c = image[y0:y0 + tile, x0:x0 + tile].float()
        if tile > NET_IN:
            c = torch.nn.functional.avg_pool2d(c[None, None], tile // NET_IN)[0, 0]
        out.append((c - c.mean()) / (c.std() + 1e-6))
    return torch.stack(out)


def peaks_of(z: torch.Tensor, zcut: float, radius: float = 5.0) -> np.ndarray:
    """(n, 2) peak positions of a z map above the cut-off (greedy 5-pixel suppression)."""
    from .analysis import find_peaks_from_zscore

    py, px = find_peaks_from_zscore(z.cpu(), zcut, radius)
    if len(py) == 0:
        return np.zeros((0, 2), dtype=int)
    return np.stack([py.numpy(), px.numpy()], 1)


def training_examples(image: torch.Tensor, best_euler: torch.Tensor, z: torch.Tensor, h: int, bins: OrientationBins,
                      zcut: float, tile: int = TILE):
    """Network inputs and targets of one searched micrograph: per tile, uniform over the
    bins of the best orientations at its z > zcut peaks, or the empty class."""
    H, W = image.shape
    tiles = tile_grid(H, W, h, tile)
    x = tile_inputs(image, tiles, tile)
    pk = peaks_of(z, zcut)
    t = torch.zeros((len(tiles), bins.num_classes))
    for i, (_, _, (oy0, oy1), (ox0, ox1)) in enumerate(tiles):
        sel = pk[(pk[:, 0] >= oy0) & (pk[:, 0] < oy1) & (pk[:, 1] >= ox0) & (pk[:, 1] < ox1)] if len(pk) else pk
        if len(sel) == 0:
            t[i, bins.empty] = 1.0
            continue
        b = bins.of(best_euler[sel[:, 0], sel[:, 1]])
        for v in b:
            t[i, int(v)] += 1.0 / len(b)
    return x, t


# ---------------------------------------------------------------------- guided search
def _tile_spectrum(real_image: torch.Tensor, y0: int, x0: int, tile: int) -> torch.Tensor:
    c = real_image[y0:y0 + tile, x0:x0 + tile].float().contiguous()
    X = torch.fft.rfft2(c)
    X[0, 0] = 0
    return X.to(torch.complex64)


def guided_search(plan_cache: dict, net: ProposalNet, bins: OrientationBins, real_image: torch.Tensor,
                  tdft: torch.Tensor, ctf: torch.Tensor, wf: torch.Tensor, euler: torch.Tensor,
                  fine_bins: np.ndarray, stats_idx: torch.Tensor, h: int, top_k: int, conf: float,
                  empty_conf: float, batch: int = 32, tile: int = TILE) -> dict:
    """Proposal-guided search of one (whitened, real-space) micrograph; ``euler`` the fine
    grid on the device, ``fine_bins`` its bins, ``stats_idx`` the orientation subset every
    searched tile runs for its per-pixel mean / std.  Returns the z map of the valid
    region (-inf where no tile searched), the tile decisions and the orientations visited.
    """
    from .backend import MatchPlan, raw_statistics

    dev = euler.device
    H, W = real_image.shape
    tiles = tile_grid(H, W, h, tile)
    x = tile_inputs(real_image, tiles, tile)
    prob = torch.softmax(net.forward(x).float(), dim=1).cpu()
    z = torch.full((H - h + 1, W - h + 1), -float("inf"))
    O = euler.shape[0]
    visited, decisions = 0, []
    key = (tile, h, batch)
    if key not in plan_cache:
        plan_cache[key] = MatchPlan(tile, tile, h, batch, 1, 1, dev)
        # every tile searches its own orientation subset (a new Euler array each time):
        # plain launches, no graph to re-capture per call
        plan_cache[key].set_option("graph", 0)
    plan = plan_cache[key]
    for i, (y0, x0, (oy0, oy1), (ox0, ox1)) in enumerate(tiles):
        p = prob[i]
        if float(p[bins.empty]) > empty_conf:
            decisions.append("empty")
            continue
        pb = p[:bins.num_bins]
        if float(pb.max()) < conf:
            idx = torch.arange(O, device=dev)
            decisions.append("exhaustive")
        else:
            top = torch.topk(pb, min(top_k, bins.num_bins)).indices.numpy()
            sel = np.nonzero(np.isin(fine_bins, top))[0]
            idx = torch.unique(torch.cat([torch.from_numpy(sel).to(dev), stats_idx]))
            decisions.append("proposal")
        plan.set_inputs(_tile_spectrum(real_image, y0, x0, tile), ctf, wf)
        # per-pixel mean / std from the statistics subset
        es = euler[stats_idx].contiguous()
        plan.reset_stats()
        plan.run(tdft, es, 0, es.shape[0])
        st = raw_statistics(plan)
        n = es.shape[0]
        mean = st["correlation_sum"] / n
        std = (st["correlation_squared_sum"] / n - mean ** 2).clamp_min(0).sqrt()
        # maximum over the visited orientations
        ev = euler[idx].contiguous()
        plan.reset_stats()
        plan.run(tdft, ev, 0, ev.shape[0])
        mip = raw_statistics(plan)["mip"].double()
        visited += int(ev.shape[0])
        zt = torch.where(std > 0, (mip - mean) / std, torch.zeros_like(mip))
        vy, vx = oy1 - oy0, ox1 - ox0
        z[oy0:oy1, ox0:ox1] = zt[oy0 - y0:oy0 - y0 + vy, ox0 - x0:ox0 - x0 + vx].float()
    return {"z": z, "decisions": decisions, "visited": visited, "tiles": len(tiles), "orientations": O}


def peak_recall(ref_peaks: np.ndarray, got_peaks: np.ndarray, radius: float = 2.0) -> float | None:
    """Fraction of reference peaks with a found peak within ``radius`` pixels."""
    if len(ref_peaks) == 0:
        return None
    if len(got_peaks) == 0:
        return 0.0
    d = np.hypot(ref_peaks[:, None, 0] - got_peaks[None, :, 0], ref_peaks[:, None, 1] - got_peaks[None, :, 1])
    return float((d.min(1) <= radius).mean())


# -------------------------------------------------------------------------- experiment
def _micrograph_inputs(template, seed: int, size: int, n_particles: int, snr: float, dev):
    """Whitened spectrum, whitening filter and whitened real-space image of one seeded
    synthetic micrograph."""
    from . import synthetic
    from .utils import preprocessing as pp

    mg = synthetic.make_micrograph(template, seed, (size, size), n_particles, 12200.0, 11800.0, 30.0, snr=snr,
                                   device=dev)
    img_dft, wf = pp.prepare_image(mg.image.to(dev, torch.float32).contiguous(), template.size)
    real = torch.fft.irfft2(img_dft, s=(size, size)) * (size * size) ** 0.5
    return img_dft, wf, real.float().contiguous(), mg


def run_experiment(dev: torch.device, n_train: int = 64, n_test: int = 16, size: int = 512, h: int = 64,
                   n_particles: int = 16, snr: float = 0.3, theta_step: float = 10.0, psi_step: float = 5.0,
                   bin_deg: float = 30.0, zcut: float = 7.5, epochs: int = 40, top_frac: float = 0.05,
                   conf: float = 0.02, empty_conf: float = 0.9, stats_frac: float = 0.05, batch: int = 32,
                   seed: int = 0, log=None, operating_points=None, tile: int = TILE, lr: float = 1e-3) -> dict:
    """Train the proposal network on the plain search of ``n_train`` seeded micrographs and
    measure the guided search on ``n_test`` held-out ones against the exhaustive search:
    recall of its z > zcut peaks, orientations visited and wall time; beside it the same
    search with ``top_k`` random bins per tile instead of the network's (the method without
    the learned part, at the same orientation budget).  ``operating_points``: further
    (confidence, empty-confidence) thresholds of the network measured the same way
    (reported under "operating_points")."""
    import time

    from . import synthetic
    from .backend import MatchPlan
    from .utils import preprocessing as pp

    def say(msg):
        if log:
            log(msg)

    t = synthetic.make_blob_template(seed, h, 500, 0.375 * h)
    vol = t.volume(device=dev)
    tdft = pp.volume_to_rfft_fourier_slice(vol).to(torch.complex64).contiguous()
    dfs = torch.tensor([0.0])
    ctf = pp.ctf_filter_stack((h, h), 12200.0, 11800.0, 30.0, dfs, torch.tensor([0.0]), device=dev)
    euler = get_uniform_euler_angles(psi_step=psi_step, theta_step=theta_step).to(dev).contiguous()
    O = euler.shape[0]
    bins = OrientationBins(bin_deg)
    fine_bins = bins.of(euler)
    rng = np.random.default_rng(seed)
    stats_idx = torch.from_numpy(np.sort(rng.choice(O, max(16, int(stats_frac * O)), replace=False))).to(dev)
    top_k = max(1, int(round(top_frac * bins.num_bins)))
    plan = MatchPlan(size, size, h, batch, 1, 1, dev)
    dvals = dfs.to(dev, torch.float32)

    def exhaustive(img_dft, wf):
        plan.set_inputs(img_dft, ctf, wf)
        plan.reset_stats()
        plan.run(tdft, euler, 0, O)
        m = plan.finalize(euler, dvals, O)
        vy = vx = size - h + 1
        z = m["scaled_mip"][:vy, :vx].cpu()
        be = torch.stack([m["best_phi"], m["best_theta"], m["best_psi"]], -1)[:vy, :vx].cpu()
        return z, be

    # training set: tiles and targets from the plain search of the training micrographs
    xs, ts = [], []
    t0 = time.perf_counter()
    for s in range(n_train):
        img_dft, wf, real, _ = _micrograph_inputs(t, 1000 + s, size, n_particles, snr, dev)
        z, be = exhaustive(img_dft, wf)
        x, tt = training_examples(real, be, z, h, bins, zcut, tile)
        xs.append(x)
        ts.append(tt)
    torch.cuda.synchronize(dev)
    gen_s = time.perf_counter() - t0
    X, T = torch.cat(xs), torch.cat(ts)
    say(f"{X.shape[0]} training tiles ({int((T[:, bins.empty] == 1).sum())} empty) in {gen_s:.1f} s")
    net = ProposalNet(bins.num_classes, 64, dev, seed=seed)
    t0 = time.perf_counter()
    hist = net.train(X, T, epochs, lr=lr, batch=32, seed=seed)
    torch.cuda.synchronize(dev)
    train_s = time.perf_counter() - t0
    say(f"training: loss {hist[0]:.3f} -> {hist[-1]:.3f} in {train_s:.1f} s")

    class RandomNet:
        """Random bin scores (the method without the network): never empty, always confident."""

        def __init__(self, seed):
            self.g = torch.Generator().manual_seed(seed)

        def forward(self, x):
            z = torch.rand((x.shape[0], bins.num_classes), generator=self.g)
            z[:, bins.empty] = -1e9
            return z * 100.0

    plans: dict = {}
    res = {"net": [], "random": [], "exhaustive_s": 0.0, "guided_s": 0.0, "random_s": 0.0}
    counts = {"net": [0, 0], "random": [0, 0]}
    decisions = {"empty": 0, "exhaustive": 0, "proposal": 0}
    hits = {"net": 0, "tiles": 0, "empty_right": 0, "empty_tiles": 0, "occupied_called_empty": 0}
    ops = [tuple(o) for o in (operating_points or [])]
    op_res = {o: {"recall": [], "visited": 0, "total": 0, "s": 0.0, "decisions": {"empty": 0, "exhaustive": 0,
                                                                                "proposal": 0}} for o in ops}
    for s in range(n_test):
        img_dft, wf, real, _ = _micrograph_inputs(t, 5000 + s, size, n_particles, snr, dev)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        z_ref, be_ref = exhaustive(img_dft, wf)
        torch.cuda.synchronize(dev)
        res["exhaustive_s"] += time.perf_counter() - t0
        ref = peaks_of(z_ref, zcut)
        # held-out tiles: is a bin of the exhaustive search's peaks among the network's top k?
        xt, tt = training_examples(real, be_ref, z_ref, h, bins, zcut, tile)
        occ = tt[:, bins.empty] < 1
        if bool(occ.any()):
            pr = torch.softmax(net.forward(xt[occ]).float(), dim=1).cpu()[:, :bins.num_bins]
            top = torch.topk(pr, min(top_k, bins.num_bins), dim=1).indices
            hits["net"] += int((torch.gather(tt[occ][:, :bins.num_bins], 1, top) > 0).any(1).sum())
            hits["tiles"] += int(occ.sum())
        hits["empty_right"] += int((torch.softmax(net.forward(xt[~occ]).float(), dim=1).cpu()[:, bins.empty]
                                    > empty_conf).sum()) if bool((~occ).any()) else 0
        hits["empty_tiles"] += int((~occ).sum())
        hits["occupied_called_empty"] += int((torch.softmax(net.forward(xt[occ]).float(), dim=1).cpu()[:, bins.empty]
                                              > empty_conf).sum()) if bool(occ.any()) else 0
        for name, model in (("net", net), ("random", RandomNet(seed + s))):
            t0 = time.perf_counter()
            g = guided_search(plans, model, bins, real, tdft, ctf, wf, euler, fine_bins, stats_idx, h, top_k,
                              conf if name == "net" else 0.0, empty_conf if name == "net" else 2.0, batch, tile)
            torch.cuda.synchronize(dev)
            res["guided_s" if name == "net" else "random_s"] += time.perf_counter() - t0
            got = peaks_of(g["z"], zcut)
            r = peak_recall(ref, got)
            if r is not None:
                res[name].append((r, len(ref)))
            counts[name][0] += g["visited"]
            counts[name][1] += g["tiles"] * O
            if name == "net":
                for d in g["decisions"]:
                    decisions[d] += 1
        for o in ops:
            t0 = time.perf_counter()
            g = guided_search(plans, net, bins, real, tdft, ctf, wf, euler, fine_bins, stats_idx, h, top_k, o[0], o[1],
                              batch, tile)
            torch.cuda.synchronize(dev)
            r = op_res[o]
            r["s"] += time.perf_counter() - t0
            rc = peak_recall(ref, peaks_of(g["z"], zcut))
            if rc is not None:
                r["recall"].append((rc, len(ref)))
            r["visited"] += g["visited"]
            r["total"] += g["tiles"] * O
            for d in g["decisions"]:
                r["decisions"][d] += 1

    def pooled(v):
        n = sum(k for _, k in v)
        return sum(r * k for r, k in v) / n if n else None

    out = {
        "train_micrographs": n_train, "test_micrographs": n_test, "image": size, "template": h,
        "orientations": O, "bins": bins.num_bins, "bin_deg": bin_deg, "top_k_bins": top_k,
        "stats_subset": int(stats_idx.numel()), "zcut": zcut, "train_tiles": int(X.shape[0]),
        "epochs": epochs, "loss_first_last": [hist[0], hist[-1]], "label_generation_s": gen_s, "train_s": train_s,
        "reference_peaks": sum(k for _, k in res["net"]),
        "recall_net": pooled(res["net"]), "recall_random_bins": pooled(res["random"]),
        "visited_fraction_net": counts["net"][0] / max(1, counts["net"][1]),
        "visited_fraction_random": counts["random"][0] / max(1, counts["random"][1]),
        "tile_decisions": decisions,
        "tile": tile,
        "held_out_top_k_hit_rate": hits["net"] / max(1, hits["tiles"]),
        "random_top_k_hit_rate": top_k / bins.num_bins,
        "held_out_occupied_tiles": hits["tiles"],
        "held_out_empty_tiles_called_empty": hits["empty_right"] / max(1, hits["empty_tiles"]),
        "held_out_occupied_tiles_called_empty": hits["occupied_called_empty"] / max(1, hits["tiles"]),
        "exhaustive_s": res["exhaustive_s"], "guided_s": res["guided_s"], "random_s": res["random_s"],
        "speedup_vs_exhaustive": res["exhaustive_s"] / max(1e-9, res["guided_s"]),
        "operating_points": [
            {"conf": o[0], "empty_conf": o[1], "recall": pooled(r["recall"]),
             "visited_fraction": r["visited"] / max(1, r["total"]), "tile_decisions": r["decisions"],
             "guided_s": r["s"], "speedup_vs_exhaustive": res["exhaustive_s"] / max(1e-9, r["s"])}
            for o, r in op_res.items()],
    }
    return out
