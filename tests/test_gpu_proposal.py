"""GPU checks of the proposal network's HIP kernels against the CPU oracle
(``oracle/proposal_net.py``, whose backward is pinned against torch autograd).

Bounds: the oracle rounds the same operands to bf16 as the kernels (weights, patch
matrices, fully connected inputs, backward output gradients) and accumulates in float32,
so the two differ by accumulation order only, plus the occasional operand that lands on
the other side of a bf16 rounding boundary: logits within 2e-3 of max |z|, every
parameter tensor's gradient within 1e-2 of its max |g| with cosine similarity above
0.9999.  The errors against the unrounded float64 network are printed."""

import pytest
import torch

from leopard_em_amd.proposal import ProposalNet

from oracle import proposal_net as pn

pytestmark = pytest.mark.gpu


def test_proposal_forward_and_gradients_match_oracle(cuda_device):
    L, N = 97, 24
    g = torch.Generator().manual_seed(3)
    params = pn.init_params(L, 3) + 0.01 * torch.randn(pn.num_params(L), generator=g)
    tiles = torch.randn((N, 64, 64), generator=g)
    t = torch.rand((N, L), generator=g, dtype=torch.float64)
    targets = (t / t.sum(1, keepdim=True)).float()
    net = ProposalNet(L, 32, cuda_device, params=params)
    z = net.forward(tiles).cpu()
    loss = net.backward(targets).cpu()
    grad = net.grads().cpu()
    zr, cache = pn.forward(params, tiles, L, bf16=True)
    lr, gr = pn.backward(cache, targets)
    z64, c64 = pn.forward(params, tiles, L, bf16=False, dtype=torch.float64)
    _, g64 = pn.backward(c64, targets.double())
    ez = float((z - zr).abs().max()) / float(zr.abs().max())
    print(f"logits: rel err vs oracle {ez:.2e}, vs float64 {float((z.double() - z64).abs().max() / z64.abs().max()):.2e}",
          flush=True)
    assert ez < 2e-3
    assert float((loss - lr).abs().max()) < 2e-3 * float(lr.abs().max())
    o = 0
    for name, s in pn.layout(L):
        k = 1
        for d in s:
            k *= d
        a, b, c = grad[o:o + k], gr[o:o + k], g64[o:o + k]
        o += k
        err = float((a - b).abs().max()) / float(b.abs().max())
        cos = float(torch.nn.functional.cosine_similarity(a.double(), b.double(), dim=0))
        cos64 = float(torch.nn.functional.cosine_similarity(a.double(), c, dim=0))
        print(f"{name}: rel err {err:.2e}, cosine {cos:.6f} (vs float64 {cos64:.6f})", flush=True)
        assert err < 1e-2 and cos > 0.9999, name


def test_proposal_training_reduces_loss(cuda_device):
    """Adam on a fixed small set: the loss falls (the network fits its training set)."""
    L, N = 20, 64
    g = torch.Generator().manual_seed(5)
    tiles = torch.randn((N, 64, 64), generator=g)
    labels = torch.randint(0, L, (N,), generator=g)
    targets = torch.nn.functional.one_hot(labels, L).float()
    net = ProposalNet(L, 64, cuda_device, seed=1)
    hist = net.train(tiles, targets, epochs=80, lr=2e-3, batch=32)
    print(f"loss: {hist[0]:.3f} -> {hist[-1]:.3f}", flush=True)
    assert hist[-1] < 0.5 * hist[0]
    acc = float((net.forward(tiles).argmax(1).cpu() == labels).float().mean())
    assert acc > 0.9
