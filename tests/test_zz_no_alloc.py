"""hush_loss_noalloc (include/hush_train.h): the training loss without a context and without allocation,
its per-workgroup partial sums in a static device array.  Through the C ABI against hush_train_loss on the
same buffers: bit for bit (the same kernels over the same workgroups), on two calls in a row, the second
with a smaller frame; sizes beyond the array are refused before any launch."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from hush_amd import _lib
    L = _lib.load()
    P, i32 = ctypes.c_void_p, ctypes.c_int32
    L.hush_loss_noalloc.argtypes = [i32, i32, i32, P, P, P, P, ctypes.c_double, P, P, P]
    L.hush_loss_noalloc.restype = i32
    L.hush_train_last_error.argtypes = [P]
    L.hush_train_last_error.restype = ctypes.c_char_p
    return L


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@pytest.mark.parametrize("fp64", [False, True], ids=["fp32", "fp64"])
def test_noalloc_loss_matches_context_loss(fp64):
    import torch
    import hush_amd
    from hush_amd import train as T
    L = _lib()
    dt = torch.float64 if fp64 else torch.float32
    for (w, h), with_prev in (((160, 96), True), ((64, 40), False)):       # the second call is smaller
        print(f"hush_loss_noalloc fp64={fp64} {w}x{h} prev={with_prev}", flush=True)
        hs = hush_amd.Hush(0, {"width": w, "height": h})
        fr = T.DiffFrame(hs, fp64=fp64)
        gen = torch.Generator(device="cuda").manual_seed(w)
        mk = lambda: torch.rand((h, w, 4), dtype=dt, device="cuda", generator=gen) * 4 - 0.5
        out, tgt = mk(), mk()
        po, pt = (mk(), mk()) if with_prev else (None, None)
        lam = 0.5 if with_prev else 0.0
        l_ref, g_ref = fr.loss(out, tgt, po, pt, lam)
        g = torch.full_like(out, float("nan"))
        l = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        r = L.hush_loss_noalloc(int(fp64), w, h, _p(out), _p(tgt), _p(po), _p(pt), lam, _p(g), _p(l), s)
        assert r == 0, L.hush_train_last_error(None)
        torch.cuda.synchronize()
        assert torch.equal(g, g_ref)
        assert l.item() == l_ref.item(), (l.item(), l_ref.item())
        fr.destroy()
        hs.destroy()


def test_noalloc_loss_size_limit():
    """4096 x 4096 (65,536 workgroups of partials) runs; 4097 x 4096 is refused with HUSH_ERR_SIZE before any
    launch (its buffers are full size all the same)"""
    import torch
    L = _lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for (w, h), want in (((4096, 4096), 0), ((4097, 4096), -3)):
        print(f"hush_loss_noalloc {w}x{h}", flush=True)
        out = torch.rand((h, w, 4), dtype=torch.float32, device="cuda")
        tgt = torch.rand((h, w, 4), dtype=torch.float32, device="cuda")
        g = torch.empty_like(out)
        l = torch.zeros((), dtype=torch.float64, device="cuda")
        r = L.hush_loss_noalloc(0, w, h, _p(out), _p(tgt), None, None, 0.0, _p(g), _p(l), s)
        torch.cuda.synchronize()
        assert r == want, (w, h, r, L.hush_train_last_error(None))
        if want == 0:
            assert torch.isfinite(l).item() and l.item() > 0
        else:
            assert b"static scratch" in L.hush_train_last_error(None)
        del out, tgt, g
