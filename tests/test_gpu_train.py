"""Gradient checks of the differentiable frame (include/hush_train.h) on the GPU.

* FP64 analytic gradients (hand-written backward kernels through TAA, recombination, the fine and mid
  a-trous levels, the temporal blend and the MLP of this and the previous frame) against FP64 central finite differences of the same forward, for all 812
  MLP weights and 200 temporal-output (A0) inputs, on 64x64 frames after three frames of history:
  beyond the difference quotient's own rounding noise (10 |loss| 2^-52 / step), relative error <= 1e-4
  for >= 99% of the checked values and <= 1e-2 for all of them (the L1 loss, the clip box and the clamps
  have kinks that a finite step can straddle).
* The FP32 backward agrees with the FP64 one to 1e-3 (relative to the largest gradient).
* Adam steps on seeded sequences lower the loss.
* The MFMA forward option (binary16 / bfloat16 MLP on the matrix cores, FP32 backward): the MLP outputs
  equal the MLP stage's MFMA kernels on the same features, the frame's output stays within the precision's
  error of the FP32 forward, and the backward of its state agrees with the FP32 backward (cosine >= 0.99
  on the pixels whose outputs agree to 1%); Adam with it lowers the loss too.
"""
import numpy as np
import pytest

import hush_amd

pytestmark = pytest.mark.gpu


def _setup(fp64, n_hist=3, kind="spheres", size=64, seed=7):
    import torch
    import hush_amd
    from hush_amd import train as T
    from hush_amd.scene import SceneGenerator
    gen = SceneGenerator(kind, size, size, seed)
    hs = hush_amd.Hush(0, {"width": size, "height": size})
    fr = T.DiffFrame(hs, fp64=fp64)
    dt = torch.float64 if fp64 else torch.float32
    w = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=dt, device="cuda")
    prev = None
    for f in range(n_hist):
        g = gen.render(f, clean=True)
        fr.prepare(T.upload(g))
        out = fr.forward(w)
        fr.commit(out)
        prev = (out, T.target_of(g, dt))
    g = gen.render(n_hist, clean=True)
    fr.prepare(T.upload(g))
    return hs, fr, w, T.target_of(g, dt), prev


def _loss(fr, w, tgt, prev, lam=0.5):
    out = fr.forward(w)
    l, g = fr.loss(out, tgt, prev[0] if prev else None, prev[1] if prev else None, lam if prev else 0.0)
    return l, g


def _rel(fd, an, noise=None):
    """|fd - an| relative to the gradient, less the finite difference's own rounding noise (`noise`:
    |loss| * 2^-52 / step per value), so gradients below the resolution of a difference quotient do not
    count as mismatches"""
    d = np.abs(fd - an)
    if noise is not None:
        d = np.maximum(d - 10.0 * noise, 0.0)
    scale = np.maximum(np.maximum(np.abs(an), np.abs(fd)), 1e-3 * np.abs(an).max())
    return d / scale


def test_fd_weights_and_inputs_fp64():
    import torch
    hs, fr, w, tgt, prev = _setup(True)
    l0, g = _loss(fr, w, tgt, prev)
    gw, ga0 = fr.backward(w, g, want_a0=True)
    an = gw.cpu().numpy()
    assert np.isfinite(an).all() and np.abs(an).max() > 0
    fd, nz = np.zeros(812), np.zeros(812)
    for k in range(812):
        eps = 1e-6 * max(1.0, abs(float(w[k])))
        nz[k] = abs(float(l0)) * 2.0 ** -52 / eps
        wp, wm = w.clone(), w.clone()
        wp[k] += eps
        wm[k] -= eps
        fd[k] = (float(_loss(fr, wp, tgt, prev)[0]) - float(_loss(fr, wm, tgt, prev)[0])) / (2 * eps)
    r = _rel(fd, an, nz)
    raw = _rel(fd, an)
    print("weights: raw rel err median", float(np.median(raw)), "max", float(raw.max()),
          "| beyond FD noise: p99", float(np.quantile(r, 0.99)), "max", float(r.max()))
    assert np.quantile(r, 0.99) <= 1e-4 and r.max() <= 1e-2, (np.argsort(-r)[:8], r[np.argsort(-r)[:8]])
    # temporal-output inputs: 200 random texel channels of A0, held fixed (no re-blend)
    fr.set_a0_fixed(True)
    a0 = fr.a0_copy()
    rng = np.random.default_rng(3)
    H, W = a0.shape[:2]
    idx = [(int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(4))) for _ in range(200)]
    ga = ga0.cpu().numpy()
    fd_a, an_a = [], []
    def fd_at(y, x, c, eps):
        ap, am = a0.clone(), a0.clone()
        ap[y, x, c] += eps
        am[y, x, c] -= eps
        fr.set_a0(ap)
        lp = float(_loss(fr, w, tgt, prev)[0])
        fr.set_a0(am)
        lm = float(_loss(fr, w, tgt, prev)[0])
        return (lp - lm) / (2 * eps)

    vals, nz_a = [], []
    for (y, x, c) in idx:
        v = float(a0[y, x, c])
        # step relative to the value: the variance channel enters through sqrt(max(gv, 1e-12))
        eps = 1e-5 * max(abs(v), 1e-3)
        fd_a.append(fd_at(y, x, c, eps))
        an_a.append(ga[y, x, c])
        nz_a.append(abs(float(l0)) * 2.0 ** -52 / eps)
        vals.append((y, x, c, v, eps))
    ra = _rel(np.array(fd_a), np.array(an_a), np.array(nz_a))
    for j in np.argsort(-ra)[:6]:
        y, x, c, v, eps = vals[j]
        print("A0 worst", (y, x, c), "value", v, "fd", fd_a[j], "an", an_a[j], "fd(eps/10)", fd_at(y, x, c, eps / 10))
    fr.set_a0(a0)
    raw = _rel(np.array(fd_a), np.array(an_a))
    print("A0: raw rel err median", float(np.median(raw)), "max", float(raw.max()),
          "| beyond FD noise: p99", float(np.quantile(ra, 0.99)), "max", float(ra.max()))
    assert np.quantile(ra, 0.99) <= 1e-4 and ra.max() <= 1e-2
    fr.destroy()
    hs.destroy()


def test_fp32_backward_matches_fp64():
    res = []
    for fp64 in (True, False):
        hs, fr, w, tgt, prev = _setup(fp64, n_hist=0)
        l, g = _loss(fr, w, tgt, prev)
        res.append((float(l), fr.backward(w, g).double().cpu().numpy()))
        fr.destroy()
        hs.destroy()
    (l64, g64), (l32, g32) = res
    assert abs(l64 - l32) <= 1e-4 * abs(l64)
    err = np.abs(g64 - g32).max() / np.abs(g64).max()
    print("fp32 vs fp64 gradient: max err / max |g|", err)
    assert err <= 1e-3


def test_fp32_backward_matches_fp64_at_512():
    """the differentiation configuration's frame size: one 512x512 'train' frame after three frames of
    history, the FP32 loss and backward against the FP64 ones (measured: loss 5e-8 relative, gradient 1.1e-4 of
    the largest)"""
    res = []
    for fp64 in (True, False):
        hs, fr, w, tgt, prev = _setup(fp64, n_hist=3, kind="train", size=512, seed=11)
        l, g = _loss(fr, w, tgt, prev)
        res.append((float(l), fr.backward(w, g).double().cpu().numpy()))
        fr.destroy()
        hs.destroy()
    (l64, g64), (l32, g32) = res
    err = np.abs(g64 - g32).max() / np.abs(g64).max()
    print("512x512: loss fp64", l64, "fp32", l32, "| fp32 vs fp64 gradient: max err / max |g|", err)
    assert abs(l64 - l32) <= 1e-6 * abs(l64)
    assert err <= 2.5e-4


def test_training_batch_at_the_differentiation_configuration():
    """fwd + bwd of the differentiation configuration's batch: 8 sequences x 16 frames of 512x512 crops of
    1920x1080 'train' scenes in FP32, every frame's gradient accumulated: the loss and the gradient are finite
    and a rerun of the same batch reproduces both bit for bit (the weight-gradient reductions run in a fixed
    order)"""
    import torch
    from hush_amd import train as T
    seq, frames, crop, scene = 8, 16, (512, 512), (1920, 1080)
    w = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=torch.float32, device="cuda",
                     requires_grad=True)
    slots = T.make_slots(seq, crop, scene)
    runs = []
    for _ in range(2):
        w.grad = None
        loss, cuts = T.run_batch(w, slots, seed=3, step=0, frames=frames, crop=crop, scene_size=scene)
        runs.append((loss, cuts, w.grad.detach().clone()))
    T.close_slots(slots)
    (l0, c0, g0), (l1, c1, g1) = runs
    print(f"8 x 16 x 512^2 batch: loss {l0:.6f}, cuts {c0}, |grad| {float(g0.norm()):.6g}")
    assert np.isfinite(l0) and torch.isfinite(g0).all() and float(g0.norm()) > 0
    assert l0 == l1 and c0 == c1 and torch.equal(g0, g1)


def test_adam_lowers_the_loss():
    from hush_amd import train as T
    w, hist = T.train(steps=12, sequences=2, frames=4, crop=(64, 64), scene_size=(160, 96), lr=3e-3, seed=1, log=None)
    losses = [l for _, l in hist]
    assert np.isfinite(losses).all()
    assert np.mean(losses[-3:]) < np.mean(losses[:3]), losses
    blob = T.to_f16_blob(w.cpu().numpy())
    assert blob.size == 812


def test_reblend_reproduces_the_temporal_output():
    """with alpha / gamma as stored (history from hush_encode frames), the differentiable frame's re-blended
    A0, rounded toward zero to binary16, is the temporal pass's stored A0 bit for bit"""
    import torch
    import hush_amd
    from hush_amd import train as T
    from hush_amd.scene import SceneGenerator
    from oracle.hush_np import Q
    from tests.parity import upload_frame
    w_, h_ = 96, 64
    gen = SceneGenerator("boxes", w_, h_, 5)
    hs = hush_amd.Hush(0, {"width": w_, "height": h_})
    surfs = None
    for f in range(3):
        surfs = upload_frame(gen.render(f), surfs)
        hs.denoise(surfs)
    fr = T.DiffFrame(hs)
    g = gen.render(3, clean=True)
    fr.prepare(T.upload(g))
    wt = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=torch.float32, device="cuda")
    fr.forward(wt)
    a0 = fr.a0_copy().double().cpu().numpy()
    ref = hs.surface("A0").download().astype(np.float64)
    assert np.array_equal(Q(True).st(a0), ref), float(np.mean(Q(True).st(a0) != ref))
    fr.destroy()
    hs.destroy()


@pytest.mark.parametrize("forward,out_tol", [("fp16-mfma", 2e-2), ("bf16-mfma", 1e-1)])
def test_mfma_forward_matches_fp32_forward(forward, out_tol):
    """one frame after three frames of history, prepared identically in two contexts, one with the FP32
    forward and one with the MFMA forward: output error p99 within out_tol relative to max(|ref|, 1e-2) and
    the loss within 2%.  The backward of the MFMA forward's state, fed the FP32 forward's loss gradient on the
    pixels whose outputs agree to 1%, matches the FP32 backward (cosine >= 0.99, norms within 10%).  The
    few pixels left out sit near TAA's inverse-tone-map singularity (1 / max(1 - lum, 1e-4)), where a
    1e-3 change of the MLP outputs moves the output by tens of percent and flips the L1 loss's signs: with
    the bf16 forward, 2 of 9216 pixels of this frame, enough to turn the raw weight gradient around.  The
    fp16 forward keeps them: its raw gradient (own loss gradient, every pixel) agrees too."""
    import torch
    from hush_amd import PRECISIONS
    ctx = []
    for f in ("fp32", forward):
        hs, fr, w, tgt, prev = _setup(False, kind="boxes", size=96, seed=9)
        assert fr._L.hush_train_set_forward_precision(fr._t, PRECISIONS[f]) == 0
        out = fr.forward(w)
        l, g = fr.loss(out, tgt, prev[0], prev[1], 0.5)
        ctx.append((hs, fr, w, out, float(l), g))
    (hs1, fr1, w, o1, l1, g1), (hs2, fr2, _, o2, l2, g2) = ctx
    rel = (o2 - o1).abs() / torch.clamp(o1.abs(), min=1e-2)
    p99 = float(torch.quantile(rel.double().flatten(), 0.99))
    stable = ((o2 - o1).abs().amax(-1) / (1 + o1.abs().amax(-1)) <= 1e-2)[..., None].to(g1.dtype)
    gm = (g1 * stable).contiguous()
    a = fr1.backward(w, gm).double().cpu().numpy()
    b = fr2.backward(w, gm).double().cpu().numpy()
    raw1 = fr1.backward(w, g1).double().cpu().numpy()
    raw2 = fr2.backward(w, g2).double().cpu().numpy()
    for hs, fr, *_ in ctx:
        fr.destroy()
        hs.destroy()
    cos = lambda x, y: float(np.dot(x, y) / (np.linalg.norm(x) * np.linalg.norm(y)))
    c, ratio, craw = cos(a, b), float(np.linalg.norm(b) / np.linalg.norm(a)), cos(raw1, raw2)
    print(forward, "output rel err p99", p99, "loss", l1, l2, "pixels outside 1%:", int((stable == 0).sum()),
          "| backward cosine", c, "norm ratio", ratio, "| raw gradient cosine", craw)
    assert p99 <= out_tol
    assert abs(l2 - l1) <= 2e-2 * abs(l1)
    assert c >= 0.99 and abs(ratio - 1.0) <= 0.1
    if forward == "fp16-mfma":
        assert craw >= 0.99


def test_mfma_forward_outputs_equal_the_mlp_stage():
    """the MFMA forward's MLP outputs (before the activations) are the MLP stage's (hush_mlp_forward) on the
    same feature rows, bit for bit: the same MFMA tile code over the same packed weights"""
    import torch
    import hush_amd
    from hush_amd import PRECISIONS
    from hush_amd.hip import DeviceBuffer, lib, check, hipMemcpyHostToDevice, hipMemcpyDeviceToHost
    hs, fr, w, tgt, prev = _setup(False, kind="boxes", size=96, seed=9)
    assert fr._L.hush_train_set_forward_precision(fr._t, PRECISIONS["fp16-mfma"]) == 0
    fr.forward(w)
    torch.cuda.synchronize()
    o = fr.mlp_outputs()
    F = np.concatenate([hs.surface(f"F{k}").download().view(np.float16) for k in range(3)], -1)
    sky = hs.surface("F3").download().view(np.float16)[..., 2].astype(np.float32) < 0
    rows = np.ascontiguousarray(F.reshape(-1, 24)[:, :20])
    n = rows.shape[0]
    fin, fout = DeviceBuffer(rows.nbytes), DeviceBuffer(n * 48)
    check(lib().hipMemcpy(fin.ptr, rows.ctypes.data, rows.nbytes, hipMemcpyHostToDevice), "h2d")
    hs.mlp_forward(fin, fout, n, "fp16-mfma")
    ref = np.empty((n, 12), np.float32)
    check(lib().hipMemcpy(ref.ctypes.data, fout.ptr, ref.nbytes, hipMemcpyDeviceToHost), "d2h")
    fr.destroy()
    hs.destroy()
    live = ~sky.reshape(-1)
    assert live.sum() > 0
    assert np.array_equal(o.reshape(-1, 12)[live], ref[live])
    assert np.all(o.reshape(-1, 12)[~live] == 0)


def test_adam_lowers_the_loss_with_mfma_forward():
    from hush_amd import train as T
    w, hist = T.train(steps=12, sequences=2, frames=4, crop=(64, 64), scene_size=(160, 96), lr=3e-3, seed=1, log=None,
                      forward="fp16-mfma")
    losses = [l for _, l in hist]
    assert np.isfinite(losses).all()
    assert np.mean(losses[-3:]) < np.mean(losses[:3]), losses


def test_prepare_refuses_bad_inputs():
    """prepare validates the G-buffer like denoise (size, presence, pitch) before anything is launched"""
    import hush_amd
    from hush_amd import train as T
    from hush_amd.hip import Surface
    from tests.parity import INPUT_KEYS
    hs = hush_amd.Hush(0, {"width": 64, "height": 48})
    fr = T.DiffFrame(hs)
    good = {k: Surface(64, 48) for k in INPUT_KEYS}
    with pytest.raises(hush_amd.HushError, match=r"^Hush\.denoise: 'motion' is 32x48, the denoiser 64x48 \(call resize\)$"):
        fr.prepare(dict(good, motion=Surface(32, 48)))
    with pytest.raises(hush_amd.HushError, match=r"^Hush\.denoise: missing input 'material'$"):
        fr.prepare({k: v for k, v in good.items() if k != "material"})
    fr.prepare(good)                                   # still usable
    fr.destroy()
    hs.destroy()


def test_resize_between_forward_and_commit_is_refused():
    """a resize after the forward leaves the denoiser's surfaces at another size and pitch: the loss and the
    commit refuse it (HUSH_ERR_SIZE) before any launch, and the resize's pending reset stands"""
    import torch
    import hush_amd
    from hush_amd import train as T
    from hush_amd.scene import SceneGenerator
    hs = hush_amd.Hush(0, {"width": 64, "height": 48})
    fr = T.DiffFrame(hs)
    w = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=torch.float32, device="cuda")
    g = SceneGenerator("spheres", 64, 48, 3).render(0, clean=True)
    fr.prepare(T.upload(g))
    out = fr.forward(w)
    torch.cuda.synchronize()
    hs.resize(32, 24)
    with pytest.raises(hush_amd.HushError, match="resized") as e:
        fr.commit(out)
    assert e.value.code == -3                          # HUSH_ERR_SIZE
    with pytest.raises(hush_amd.HushError, match="resized"):
        fr.loss(out, T.target_of(g, torch.float32))
    assert hs.frame == 0
    fr.destroy()
    hs.destroy()


def test_training_step_at_the_end_to_end_configuration():
    """The end-to-end training configuration's step: 16 sequences x 24 frames of 256x256 crops of 1920x1080
    'train' scenes (300 primitives, 30% moving, a camera cut every 48 frames resets the history), FP16-MFMA
    forward and FP32 backward.  The batch holds camera cuts; its loss is finite; the gradient points the way of
    the FP32-forward gradient of the same batch (cosine >= 0.99); one Adam step lowers the batch's loss."""
    import torch
    from hush_amd import train as T
    seq, frames, crop, scene = 16, 24, (256, 256), (1920, 1080)
    w = torch.tensor(T.random_weights(0), dtype=torch.float32, device="cuda", requires_grad=True)
    res = {}
    for fwd in ("fp32", "fp16-mfma"):
        slots = T.make_slots(seq, crop, scene, forward=fwd)
        w.grad = None
        loss, cuts = T.run_batch(w, slots, seed=0, step=0, frames=frames, crop=crop, scene_size=scene)
        res[fwd] = (loss, cuts, w.grad.detach().clone())
        if fwd == "fp32":
            T.close_slots(slots)
    l16, cuts, g16 = res["fp16-mfma"]
    l32, _, g32 = res["fp32"]
    cos = float(torch.dot(g16, g32) / (g16.norm() * g32.norm()))
    print(f"loss fp16-mfma {l16:.6f} fp32 {l32:.6f}, cuts {cuts}, gradient cosine {cos:.5f}")
    assert cuts >= 1
    assert np.isfinite(l16) and np.isfinite(l32) and torch.isfinite(g16).all()
    assert abs(l16 - l32) <= 1e-2 * l32, (l16, l32)
    assert cos >= 0.99, cos
    opt = torch.optim.Adam([w], lr=1e-3)
    w.grad = g16
    opt.step()
    after, _ = T.run_batch(w, slots, seed=0, step=0, frames=frames, crop=crop, scene_size=scene, backward=False)
    T.close_slots(slots)
    print(f"batch loss before {l16:.6f} after one Adam step {after:.6f}")
    assert np.isfinite(after) and after < l16, (l16, after)
