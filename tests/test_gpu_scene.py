"""The GPU render of the procedural scenes (include/hush_scene.h) against the CPU generator
(hush_amd/csrc/scene.c) on the same frames and crops: normals / depth, motion / ids, material, emission and
the clean target bit-exact; radiance bit-exact except where the 1 spp noise's log / cos / exp (device
library vs C library, within an FP64 ulp) lands a binary16 rounding differently — at most 1 binary16 ulp
and at most 1e-4 of the values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("radiance", "normalDepth", "motion", "material", "emission")


@pytest.mark.parametrize("kind,W,H,seed,frames,crop", [
    ("boxes", 320, 180, 0, (0, 5), None),
    ("spheres", 256, 144, 1, (3,), None),
    ("train", 640, 360, 3, (47, 48, 49), (100, 40, 200, 150)),
])
def test_gpu_render_matches_cpu(kind, W, H, seed, frames, crop):
    import torch
    from hush_amd.scene import GpuSceneGenerator
    g = GpuSceneGenerator(kind, W, H, seed)
    x0, y0, w, h = crop if crop else (0, 0, W, H)
    for f in frames:
        ref = g.cpu.render(f, crop=crop, clean=True)
        out = {k: torch.empty((h, w, 4), dtype=torch.float16, device="cuda") for k in KEYS}
        clean = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        g.render(f, out, crop=crop, clean=clean)
        torch.cuda.synchronize()
        for k in KEYS:
            a = out[k].cpu().numpy().view(np.uint16)
            b = ref[k]
            if k == "radiance":
                diff = a != b
                assert diff.mean() <= 1e-4, (f, k, float(diff.mean()))
                d = np.abs(a.astype(np.int32) - b.astype(np.int32))
                assert d.max() <= 1, (f, k, int(d.max()))
            else:
                assert np.array_equal(a, b), (f, k, float(np.mean(a != b)))
        assert np.array_equal(clean.cpu().numpy(), ref["clean"]), f
    g.close()
