"""Host-side logic without a GPU: scene generators, the enclosed-pocket
analysis, the kernel launch shim's dimension checks and the bench's
one-GPU contract."""
import os
import subprocess
import sys

import numpy as np
import pytest

from flub_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("make", [lambda: scenes.dam_break(16, 0), lambda: scenes.dam_break_spheres(16, 1),
                                  lambda: scenes.flow_past_sphere(16, 2), lambda: scenes.random_pool(16, 1000)])
def test_scene_layout_and_determinism(make):
    a, b = make(), make()
    n = a["n"]
    assert a["tag"].dtype == np.int32 and set(np.unique(a["tag"])) <= {0, 1, 2}
    for c, shp in (("u", (n + 1, n, n)), ("v", (n, n + 1, n)), ("w", (n, n, n + 1))):
        assert a[c].shape == shp and a[c].dtype == np.float32
        assert np.array_equal(a[c], b[c])
    assert np.all(a["u"][0] == 0) and np.all(a["u"][-1] == 0)
    assert np.all(a["v"][:, 0] == 0) and np.all(a["w"][:, :, -1] == 0)
    mu, _, _ = scenes.fluid_face_masks(a["tag"])
    assert np.all(a["u"][~mu] == 0)


def test_dam_break_geometry():
    s = scenes.dam_break(20, 0)
    c = scenes.centres(20)
    exp = (c[:, None, None] < 0.4) & (c[None, :, None] < 0.6)
    assert np.array_equal(s["tag"] == 1, np.broadcast_to(exp, (20, 20, 20)))
    assert not np.any(s["tag"] == 2)


def test_flow_past_sphere_boundaries():
    s = scenes.flow_past_sphere(32, 2)
    t = s["tag"]
    assert np.all(t[0] == 2) and np.all(t[-1] == 0)
    assert t[16, 16, 16] == 2 and t[8, 8, 8] == 1
    assert np.all(s["ur"][1] == 1.0)


def test_analyze_retags_enclosed_pocket():
    from flub_amd.multigrid.mgpcg import MGPCG

    tag = np.zeros((8, 8, 8), np.int32)
    tag[:, :, :] = 1
    tag[:, :, 4:] = 0                                  # open fluid touching air
    tag[1:4, 1:4, 5] = 2
    tag[1:4, 1:4, 7] = 2
    tag[1:4, 1, 6] = 2; tag[1:4, 3, 6] = 2; tag[1, 1:4, 6] = 2; tag[3, 1:4, 6] = 2
    tag[2, 2, 6] = 1                                   # 1-cell pocket: not retagged (len must be > 1)
    tag[5:7, 5:7, 0:2] = 1
    out = MGPCG.analyze(None, tag)
    assert out[2, 2, 6] == 1
    # a 2-cell enclosed fluid pocket with no AIR is retagged SOLID
    pocket = np.full((6, 6, 6), 2, np.int32)
    pocket[2, 2, 2] = 1
    pocket[2, 2, 3] = 1
    out2 = MGPCG.analyze(None, pocket)
    assert out2[2, 2, 2] == 2 and out2[2, 2, 3] == 2


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_div_small_is_ieee_division(tmp_path):
    """The blocked smoother's reciprocal division (csrc/div_small.h) equals
    x / d for d = 1..6 on every x near its FMA bound and on a sample of the
    rest (tests/div_small_check.cpp, built for the host CPU)."""
    exe = tmp_path / "div_small_check"
    src = os.path.join(ROOT, "tests", "div_small_check.cpp")
    subprocess.run(["g++", "-O2", "-fopenmp", "-std=c++17", "-ffp-contract=off", "-o", str(exe), src, "-lm"],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env={**os.environ, "OMP_NUM_THREADS": "8"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout


def test_oracle_under_address_and_ub_sanitizers():
    """The CPU oracle built with -fsanitize=address,undefined (make -C oracle
    sanitize) runs its driver over every routine family on cubic and
    non-cubic grids without a report."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       env={**os.environ, "OMP_NUM_THREADS": "2"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized oracle run complete" in r.stdout
