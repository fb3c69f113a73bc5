"""GPU test of the statistics merge (tdtm_stats_merge), the per-device aggregation rule of
``backend/process_results.py:7-64``: the larger MIP wins with the first operand kept on
ties (numpy argmax over the device axis, ``:38-44``), sums and sums of squares add."""

import pytest
import torch

from leopard_em_amd import _lib
from leopard_em_amd.backend import MatchPlan, raw_statistics

from test_gpu_match import _inputs

pytestmark = pytest.mark.gpu


def test_two_plans_merged_equal_one_plan(cuda_device):
    """Two plans over disjoint orientation halves, merged, against one plan over all."""
    print("merge: two halves vs one plan, 512x512 h64 D2 O16 B4", flush=True)
    H = W = 512
    h, D, O, B = 64, 2, 16, 4
    img_dft, tdft, ctf, wf, dfs, pix, eul = _inputs(H, W, h, D, 1, O, seed=31)
    tdft_d, eul_d = tdft.to(cuda_device), eul.to(cuda_device)
    plans = []
    for lo, hi in ((0, O), (0, O // 2), (O // 2, O)):
        p = MatchPlan(H, W, h, B, D, 1, cuda_device)
        p.set_inputs(img_dft, ctf, wf)
        p.reset_stats()
        p.run(tdft_d, eul_d, lo, hi)
        plans.append(p)
    whole, first, second = plans
    first.merge_from(second)
    torch.cuda.synchronize()
    a, b = raw_statistics(whole), raw_statistics(first)
    assert torch.equal(a["mip"], b["mip"])
    assert torch.equal(a["best_global_index"], b["best_global_index"])
    for k in ("correlation_sum", "correlation_squared_sum"):
        scale = float(a[k].abs().max())
        err = float((a[k] - b[k]).abs().max()) / scale
        print(f"{k}: merged vs single-plan rel err {err:.2e}", flush=True)
        assert err < 1e-6


def test_merge_rule_ties_and_sums(cuda_device):
    """Element-wise rule on constructed maps: strict > takes b, ties keep a; (hi, lo) sums
    add to within float32 rounding of the float64 sum."""
    n = 4096
    g = torch.Generator().manual_seed(0)
    mip_a = torch.randn(n, generator=g)
    mip_b = mip_a.clone()
    sel = torch.arange(n)
    mip_b[sel % 3 == 1] += 0.5    # b larger: b wins
    mip_b[sel % 3 == 2] -= 0.5    # b smaller: a stays
    # sel % 3 == 0: exact ties, a must stay
    idx_a = torch.arange(n, dtype=torch.int32)
    idx_b = idx_a + 100000
    s = [torch.randn(n, generator=g) * 1e3 for _ in range(4)]
    t = [torch.randn(n, generator=g) * 1e3 for _ in range(4)]
    s[1] *= 1e-7  # lo parts are small against hi parts
    s[3] *= 1e-7
    t[1] *= 1e-7
    t[3] *= 1e-7
    dev = [x.to(cuda_device).contiguous() for x in (mip_a, idx_a, *s)]
    devb = [x.to(cuda_device).contiguous() for x in (mip_b, idx_b, *t)]
    lib = _lib.load()
    st = _lib.c_void_p(torch.cuda.current_stream(cuda_device).cuda_stream)
    _lib.check(lib.tdtm_stats_merge(n, *[_lib.ptr(x) for x in dev], *[_lib.ptr(x) for x in devb], st))
    torch.cuda.synchronize()
    mip, idx = dev[0].cpu(), dev[1].cpu()
    b_wins = sel % 3 == 1
    assert torch.equal(mip, torch.where(b_wins, mip_b, mip_a))
    assert torch.equal(idx, torch.where(b_wins, idx_b, idx_a))
    for j in (0, 2):
        got = dev[2 + j].cpu().double() + dev[3 + j].cpu().double()
        ref = s[j].double() + s[j + 1].double() + t[j].double() + t[j + 1].double()
        assert float((got - ref).abs().max()) <= 1e-6 * float(ref.abs().max())


def test_merge_rejects_null_pointers(cuda_device):
    lib = _lib.load()
    x = torch.zeros(16, device=cuda_device)
    ptrs = [_lib.ptr(x)] * 12
    ptrs[7] = _lib.c_void_p(0)
    assert lib.tdtm_stats_merge(16, *ptrs, None) != 0
    assert b"null" in lib.tdtm_last_error()
