"""Pin of the oracle's proposal-network restatement (CPU): its hand-written backward
against torch autograd of the same network built from torch.nn.functional (float64,
rounding off), and the bf16 rounding points against the unrounded float64 result."""

import torch

from oracle import proposal_net as pn


def _case(L=37, N=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    params = pn.init_params(L, seed) + 0.01 * torch.randn(pn.num_params(L), generator=g)
    tiles = torch.randn((N, 64, 64), generator=g)
    t = torch.rand((N, L), generator=g, dtype=torch.float64)
    # rows sum to one in float64: the loss gradient softmax - t assumes it
    return params, tiles, t / t.sum(1, keepdim=True)


def test_manual_backward_equals_autograd_float64():
    L = 37
    params, tiles, targets = _case(L)
    z, cache = pn.forward(params, tiles, L, bf16=False, dtype=torch.float64)
    loss, grad = pn.backward(cache, targets)
    loss_ref, grad_ref = pn.autograd_reference(params, tiles, targets, L)
    assert torch.allclose(loss, loss_ref, rtol=1e-12, atol=1e-12)
    assert float((grad - grad_ref).abs().max()) <= 1e-10 * float(grad_ref.abs().max())
    assert pn.num_params(L) == grad.numel()


def test_im2col_col2im_are_transposes():
    g = torch.Generator().manual_seed(1)
    x = torch.randn((2, 8, 8, 3), generator=g, dtype=torch.float64)
    d = torch.randn((2 * 8 * 8, 27), generator=g, dtype=torch.float64)
    # <im2col(x), d> = <x, col2im(d)>
    lhs = float((pn.im2col(x) * d).sum())
    rhs = float((x * pn.col2im(d, 2, 8, 8, 3)).sum())
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_bf16_rounding_is_a_small_perturbation():
    L = 37
    params, tiles, targets = _case(L, seed=2)
    z16, c16 = pn.forward(params, tiles, L, bf16=True)
    z64, c64 = pn.forward(params, tiles, L, bf16=False, dtype=torch.float64)
    assert float((z16.double() - z64).abs().max()) < 0.05 * float(z64.abs().max())
    _, g16 = pn.backward(c16, targets)
    _, g64 = pn.backward(c64, targets)
    cos = float(torch.nn.functional.cosine_similarity(g16.double(), g64, dim=0))
    assert cos > 0.999
