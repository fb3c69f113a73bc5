"""tdtm_prepare_image_noalloc (scratch-array variant of tdtm_prepare_image) against the
allocating entry point, through the C ABI: two calls in a row, the second on a smaller
image, each compared with tdtm_prepare_image on the same input.  Both run the same
kernels on the same twiddle values, so the outputs are expected to agree bit for bit.
Oversized requests are refused with an error."""

import pytest
import torch

from leopard_em_amd import _lib

pytestmark = pytest.mark.gpu


def _allocating(lib, img, h, whitening, bins, fmax, st):
    H, W = img.shape
    nb = lib.tdtm_prepare_image_workspace_bytes(H, W, bins)
    assert nb > 0
    ws = torch.empty(nb, dtype=torch.uint8, device=img.device)
    X = torch.empty((H, W // 2 + 1), dtype=torch.complex64, device=img.device)
    wt = torch.empty((h, h // 2 + 1), dtype=torch.float32, device=img.device)
    _lib.check(lib.tdtm_prepare_image(_lib.ptr(img), H, W, h, whitening, bins, fmax, _lib.ptr(X), _lib.ptr(wt),
                                      _lib.ptr(ws), st))
    return X, wt


def _noalloc(lib, img, h, whitening, bins, fmax, st):
    H, W = img.shape
    X = torch.empty((H, W // 2 + 1), dtype=torch.complex64, device=img.device)
    wt = torch.empty((h, h // 2 + 1), dtype=torch.float32, device=img.device)
    _lib.check(lib.tdtm_prepare_image_noalloc(_lib.ptr(img), H, W, h, whitening, bins, fmax, _lib.ptr(X),
                                              _lib.ptr(wt), st))
    return X, wt


def test_noalloc_matches_allocating_entry_point(cuda_device):
    lib = _lib.load()
    st = _lib.c_void_p(torch.cuda.current_stream(cuda_device).cuda_stream)
    g = torch.Generator(device=cuda_device).manual_seed(0)
    # first call on a large image, then a smaller one with other filter settings: the
    # second must not see anything the first left in the scratch array
    for (H, W, h, whitening, bins, fmax) in ((2048, 4096, 256, 1, 0, 0.5), (256, 512, 64, 1, 100, 0.4),
                                             (128, 128, 32, 0, 0, 0.5)):
        print(f"noalloc vs allocating: {H}x{W} h{h} whitening={whitening} bins={bins}", flush=True)
        img = torch.randn((H, W), generator=g, device=cuda_device)
        X0, w0 = _allocating(lib, img, h, whitening, bins, fmax, st)
        X1, w1 = _noalloc(lib, img, h, whitening, bins, fmax, st)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(X0), torch.view_as_real(X1))
        assert torch.equal(w0, w1)


def test_noalloc_refuses_oversized_requests(cuda_device):
    lib = _lib.load()
    img = torch.zeros((64, 96), device=cuda_device)
    X = torch.empty((64, 49), dtype=torch.complex64, device=cuda_device)
    wt = torch.empty((32, 17), device=cuda_device)
    assert lib.tdtm_prepare_image_noalloc(_lib.ptr(img), 64, 96, 32, 1, 0, 0.5, _lib.ptr(X), _lib.ptr(wt), None) != 0
    assert b"powers of two" in lib.tdtm_last_error()
    img = torch.zeros((64, 64), device=cuda_device)
    assert lib.tdtm_prepare_image_noalloc(_lib.ptr(img), 64, 64, 32, 1, 9000, 0.5, _lib.ptr(X), _lib.ptr(wt),
                                          None) != 0
    assert b"8192 frequency bins" in lib.tdtm_last_error()
