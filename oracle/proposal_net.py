"""CPU restatement of the orientation-proposal network (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference; this module restates the HIP network of
``leopard-em_amd/csrc/proposal.hip`` in torch on the CPU so that its forward pass and its
gradients can be checked:
* layers: four 3x3 conv (padding 1) + ReLU + 2x2 average pooling (16, 32, 64, 64
  channels), fc 1024 -> 256 + ReLU, fc 256 -> L; NHWC activations;
* parameter vector layout: per conv layer W (Cout, 3, 3, Cin), b (Cout); W_fc1 (256, 1024)
  with the 4 x 4 x 64 features in (y, x, c) order, b_fc1; W_fc2 (L, 256), b_fc2;
* rounding (``bf16=True``): the operands of every matrix product are rounded to bf16
  (round to nearest even) -- weights, patch matrices, fully connected inputs and the
  output gradients of the backward products -- with float32 accumulation, as the HIP
  kernels do;
* loss: softmax cross-entropy against target distributions, averaged over the batch;
* ``backward`` is written out by hand (patch matrices, col2im, pooling and ReLU masks).
  ``tests/test_oracle_proposal.py`` pins it against torch autograd of the same network
  built from ``torch.nn.functional.conv2d`` / ``avg_pool2d`` / ``linear`` in float64 with
  the rounding off.
"""

from __future__ import annotations

import torch

CH = (1, 16, 32, 64, 64)
HID = 256
FEAT = 4 * 4 * 64
EDGE = 64


def layout(L: int):
    """[(name, shape)] of the parameter vector in order."""
    out = []
    for i in range(4):
        out += [(f"conv{i}.w", (CH[i + 1], 3, 3, CH[i])), (f"conv{i}.b", (CH[i + 1],))]
    out += [("fc1.w", (HID, FEAT)), ("fc1.b", (HID,)), ("fc2.w", (L, HID)), ("fc2.b", (L,))]
    return out


def num_params(L: int) -> int:
    n = 0
    for _, s in layout(L):
        k = 1
        for d in s:
            k *= d
        n += k
    return n


def unflatten(p: torch.Tensor, L: int) -> dict:
    out, o = {}, 0
    for name, s in layout(L):
        k = 1
        for d in s:
            k *= d
        out[name] = p[o:o + k].reshape(s)
        o += k
    return out


def _r(x: torch.Tensor, bf16: bool) -> torch.Tensor:
    return x.to(torch.bfloat16).to(x.dtype) if bf16 else x


def im2col(x: torch.Tensor) -> torch.Tensor:
    """(N, H, W, C) -> (N H W, 9 C), column (ky, kx, c), zero padding 1."""
    N, H, W, C = x.shape
    xp = torch.nn.functional.pad(x, (0, 0, 1, 1, 1, 1))
    taps = [xp[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)]
    return torch.stack(taps, 3).reshape(N * H * W, 9 * C)


def col2im(dcols: torch.Tensor, N: int, H: int, W: int, C: int) -> torch.Tensor:
    """Transpose of im2col: (N H W, 9 C) -> (N, H, W, C)."""
    d = dcols.reshape(N, H, W, 9, C)
    out = torch.zeros((N, H + 2, W + 2, C), dtype=dcols.dtype)
    for t in range(9):
        ky, kx = divmod(t, 3)
        out[:, ky:ky + H, kx:kx + W, :] += d[:, :, :, t, :]
    return out[:, 1:H + 1, 1:W + 1, :]


def avgpool(x: torch.Tensor) -> torch.Tensor:
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(dim=(2, 4))


def avgpool_grad(d: torch.Tensor) -> torch.Tensor:
    return 0.25 * d.repeat_interleave(2, 1).repeat_interleave(2, 2)


def forward(params: torch.Tensor, tiles: torch.Tensor, L: int, bf16: bool = True, dtype=torch.float32):
    """Logits (N, L) and the cache ``backward`` needs."""
    P = {k: v.to(dtype) for k, v in unflatten(params, L).items()}
    x = tiles.to(dtype).reshape(-1, EDGE, EDGE, 1)
    cache = {"inputs": [], "conv": [], "cols": [], "P": P, "bf16": bf16}
    for i in range(4):
        N, H, W, C = x.shape
        cols = _r(im2col(x), bf16)
        w = _r(P[f"conv{i}.w"].reshape(CH[i + 1], -1), bf16)
        y = torch.relu(cols @ w.t() + P[f"conv{i}.b"])
        cache["inputs"].append(x)
        cache["cols"].append(cols)
        cache["conv"].append(y.reshape(N, H, W, CH[i + 1]))
        x = avgpool(cache["conv"][-1])
    F = x.reshape(x.shape[0], FEAT)
    h1 = torch.relu(_r(F, bf16) @ _r(P["fc1.w"], bf16).t() + P["fc1.b"])
    z = _r(h1, bf16) @ _r(P["fc2.w"], bf16).t() + P["fc2.b"]
    cache.update(F=F, h1=h1, z=z)
    return z, cache


def backward(cache: dict, targets: torch.Tensor):
    """(per-sample loss, flat gradient of the batch-mean loss) by hand."""
    bf16, P, z = cache["bf16"], cache["P"], cache["z"]
    t = targets.to(z.dtype)
    N = z.shape[0]
    lse = torch.logsumexp(z, dim=1, keepdim=True)
    loss = -(t * (z - lse)).sum(1)
    dz = (torch.exp(z - lse) - t) / N
    g = {}
    dzr = _r(dz, bf16)
    g["fc2.w"] = dzr.t() @ _r(cache["h1"], bf16)
    g["fc2.b"] = dz.sum(0)
    dh1 = dzr @ _r(P["fc2.w"], bf16)
    dh1 = dh1 * (cache["h1"] > 0)
    dh1r = _r(dh1, bf16)
    g["fc1.w"] = dh1r.t() @ _r(cache["F"], bf16)
    g["fc1.b"] = dh1.sum(0)
    dP = (dh1r @ _r(P["fc1.w"], bf16)).reshape(N, 4, 4, 64)
    for i in range(3, -1, -1):
        y = cache["conv"][i]
        Nb, H, W, Co = y.shape
        dY = (avgpool_grad(dP) * (y > 0)).reshape(-1, Co)
        g[f"conv{i}.b"] = dY.sum(0)
        dYr = _r(dY, bf16)
        g[f"conv{i}.w"] = (dYr.t() @ cache["cols"][i]).reshape(Co, 3, 3, CH[i])
        if i > 0:
            w = _r(P[f"conv{i}.w"].reshape(Co, -1), bf16)
            dP = col2im(dYr @ w, Nb, H, W, CH[i])
    L = z.shape[1]
    flat = torch.cat([g[name].reshape(-1) for name, _ in layout(L)])
    return loss, flat


def autograd_reference(params: torch.Tensor, tiles: torch.Tensor, targets: torch.Tensor, L: int):
    """The same network from torch.nn.functional in float64 (no rounding), with autograd:
    (per-sample loss, flat gradient).  Used only to pin ``backward``."""
    p = params.detach().to(torch.float64).clone().requires_grad_(True)
    P = unflatten(p, L)
    x = tiles.to(torch.float64).reshape(-1, 1, EDGE, EDGE)
    for i in range(4):
        w = P[f"conv{i}.w"].permute(0, 3, 1, 2)  # (Co, Ci, 3, 3)
        x = torch.nn.functional.avg_pool2d(torch.relu(torch.nn.functional.conv2d(x, w, P[f"conv{i}.b"], padding=1)), 2)
    F = x.permute(0, 2, 3, 1).reshape(x.shape[0], FEAT)  # (y, x, c) order
    h1 = torch.relu(torch.nn.functional.linear(F, P["fc1.w"], P["fc1.b"]))
    z = torch.nn.functional.linear(h1, P["fc2.w"], P["fc2.b"])
    lse = torch.logsumexp(z, 1, keepdim=True)
    loss = -(targets.to(torch.float64) * (z - lse)).sum(1)
    loss.mean().backward()
    return loss.detach(), p.grad.detach()


def init_params(L: int, seed: int = 0) -> torch.Tensor:
    """He-normal weights, zero biases (float32)."""
    g = torch.Generator().manual_seed(seed)
    parts = []
    for name, s in layout(L):
        if name.endswith(".b"):
            parts.append(torch.zeros(s))
        else:
            fan_in = 1
            for d in s[1:]:
                fan_in *= d
            parts.append(torch.randn(s, generator=g) * (2.0 / fan_in) ** 0.5)
    return torch.cat([t.reshape(-1) for t in parts]).float()
