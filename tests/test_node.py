"""The Node.js host side: the JavaScript Hush class over the N-API addon and the C ABI."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "hush_amd", "node", "hush_napi.node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is missing")


@needs_node
def test_api_errors_without_gpu():
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "api_errors.js")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "api_errors: ok" in r.stdout


@needs_node
@pytest.mark.gpu
def test_node_sequence_matches_oracle(tmp_path):
    import hush_amd
    from hush_amd.scene import SceneGenerator
    from oracle.ho import Oracle
    from tests.parity import compare
    w, h, frames = 160, 96, 4
    gen = SceneGenerator("boxes", w, h, seed=11)
    o = Oracle(w, h, hush_amd.default_weights())
    refs = []
    for f in range(frames):
        g = gen.render(f)
        for k in ("radiance", "normalDepth", "motion", "material", "emission"):
            g[k].tofile(tmp_path / f"f{f}_{k}.bin")
        refs.append(o.denoise(g))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "denoise_sequence.js"), str(tmp_path), str(w), str(h),
                        str(frames)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    for f in range(frames):
        out = np.fromfile(tmp_path / f"out{f}.bin", np.float16).reshape(h, w, 4).astype(np.float32)
        st = compare(out[..., :3], refs[f][..., :3])
        assert st["frac_rel_gt_1e-3"] <= 1e-3 and st["max_rel"] <= 1e-2, (f, st)
        assert np.all(out[..., 3] == 1.0)
    timing = json.loads((tmp_path / "timing.json").read_text())
    assert [p["label"] for p in timing] == list(hush_amd.PASS_LABELS)
    assert all(p["ms"] > 0 for p in timing)


@needs_node
@pytest.mark.gpu
def test_node_svgf_sequence_matches_oracle(tmp_path):
    from hush_amd.scene import SceneGenerator
    from oracle.ho import SvgfOracle
    from tests.parity import compare
    w, h, frames = 128, 80, 4
    gen = SceneGenerator("boxes", w, h, seed=13)
    o = SvgfOracle(w, h)
    refs, prev = [], None
    for f in range(frames):
        g = gen.render(f)
        for k in ("radiance", "normalDepth", "motion", "material", "emission"):
            g[k].tofile(tmp_path / f"f{f}_{k}.bin")
        src = prev if prev is not None else g
        refs.append(o.encode(dict(g, prevNormalDepth=src["normalDepth"], prevMotion=src["motion"])))
        prev = g
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "svgf_sequence.js"), str(tmp_path), str(w), str(h),
                        str(frames)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    for f in range(frames):
        out = np.fromfile(tmp_path / f"out{f}.bin", np.float16).reshape(h, w, 4).astype(np.float32)
        st = compare(out, refs[f])
        assert st["frac_rel_gt_1e-3"] <= 1e-3 and st["max_rel"] <= 1e-2, (f, st)
    timing = json.loads((tmp_path / "timing.json").read_text())
    assert [p["label"] for p in timing] == ["temporal", "variance"] + ["atrous"] * 5 + ["resolve"]
