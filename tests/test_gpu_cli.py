"""The `pt` command-line renderer (the reference's main.cpp flow over the C++ facade)."""
import os
import subprocess

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import imageio

pytestmark = pytest.mark.gpu

PT = os.path.join(os.path.dirname(pt.__file__), "pt")


@pytest.mark.skipif(not os.path.exists(PT), reason="pt CLI not built")
def test_cli_renders_cornell_with_both_integrators(tmp_path):
    imgs = {}
    for integ in ("path", "wavefront"):
        out = tmp_path / f"cornell_{integ}.json"
        scene = open(os.path.join(os.path.dirname(__file__), "..", "scenes", "cornell_box.json")).read()
        scene = scene.replace('"cornell_box.hdr"', f'"{tmp_path / ("out_" + integ + ".hdr")}"')
        out.write_text(scene)
        r = subprocess.run([PT, str(out), "--spp", "16", "--integrator", integ], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        assert "Image saved" in r.stdout
        imgs[integ] = imageio.read_hdr(str(tmp_path / ("out_" + integ + ".hdr")))
    a, b = imgs["path"], imgs["wavefront"]
    assert a.shape == b.shape and a.shape[0] == 512
    # the same random numbers: the same image up to RGBE quantisation
    assert np.allclose(a, b, rtol=1e-2, atol=1e-3)
    assert 0.05 < a[..., :3].mean() < 5.0


@pytest.mark.skipif(not os.path.exists(PT), reason="pt CLI not built")
def test_cli_rejects_missing_scene(tmp_path):
    r = subprocess.run([PT, str(tmp_path / "nope.json")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
