"""The Node host (hahenn-_amd/js over the N-API addon): CPU logic checks, and on the GPU a preset run
through FluidSim replayed by the oracle from the recorded boundary inputs."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_host_cpu_logic():
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "host_cpu.js")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_addon_loads():
    r = subprocess.run([NODE, "-e", "const a=require(process.argv[1]); if (a.version()!==2) process.exit(3); "
                        "const c=a.configDefault('high'); if (c.iters!==56 || c.cap!==700000) process.exit(4);",
                        os.path.join(ROOT, "hahenn-_amd", "suiri_napi.node")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_fluidsim_dam_preset_matches_oracle(tmp_path):
    import oracle as O
    frames = 30
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "run_scene.js"), "dam", str(frames), str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert '"statsAsync":"ok"' in r.stdout   # the Promise-based stats agree with the blocking ones
    tr = json.load(open(tmp_path / "trace.json"))
    opts = tr["opts"]
    q = O.quality("low", ripple_R=opts["ripple_R"], cap=opts["cap"], ww_cap=opts["ww_cap"])
    o = O.Oracle((tr["tank"]["w"], tr["tank"]["h"], tr["tank"]["d"]), q, threads=8)
    rows = np.array(tr["rows"], np.float32).reshape(-1, 28)
    o.set_objects(rows)
    o.emit_now(tr["emits"])
    for f in range(frames):
        subs = tr["substeps"][f]
        n = len(subs)
        assert o.L.or_substep_count(o.ptr, 1 / 120) == n
        o.L.or_step_begin(o.ptr)
        for recs in subs:
            o.step_substep(1 / 120 / n, recs)
        o.L.or_frame_tail(o.ptr, 1 / 120)
        o.L.or_poll_stats(o.ptr)
    cnt = int(o.info()["count"])
    assert tr["stats"]["count"] == cnt
    gp = np.fromfile(tmp_path / "particles.bin", np.float32).reshape(-1, 8)[:cnt]
    op = o.field("parts")[:cnt]
    scale = np.abs(op[:, :3]).max()
    assert np.abs(gp[:, :3] - op[:, :3]).max() <= 1e-5 * scale
    assert np.array_equal(gp[:, 3], op[:, 3])
    pres = np.fromfile(tmp_path / "pres.bin", np.float32)
    assert np.abs(pres - o.field("pres")).max() <= 1e-5 * max(np.abs(o.field("pres")).max(), 1e-30)


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_bake_server_job_equals_plain_fluidsim(tmp_path):
    """A faucet preset served by the Node BakeServer beside another job equals the same preset run
    alone through FluidSim with one substep per frame, particle for particle; its bake frames land
    in the requested directory."""
    frames = 12
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "serve_scene.js"), "faucet", str(frames), str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["plainCount"] > 0 and out["frames"] == frames
    assert out["equal"], out
    assert sorted(out["done"]) == ["done", "done"]
    names = sorted(os.listdir(tmp_path))
    assert "frame_00000.npz" in names and "frame_%05d.npz" % frames in names


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_converged_pressure_solver():
    """FluidSim.setPressureSolver('converged') through the addon: every frame's solve ends at the
    tolerance without breakdown, and an unknown solver name is an error."""
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "converged_scene.js"), "6"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert "unknown pressure solver" in out["unknown"]
    for f in out["frames"]:
        assert not f["breakdown"] and 1 <= f["iters"] <= 40 and f["residPerS"] <= 1e-5 < f["resid0PerS"], f
    assert out["frames"][-1]["count"] == out["frames"][0]["count"] > 0
