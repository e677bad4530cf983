"""The droplet layer on the GPU: sub-grid droplets (lognormal diameters, uniform speeds and impact
points) deposit 2 m v impulses into the ripple source and splash whitewater when K = We Oh^-0.4
exceeds the jet threshold (60 here, the reference's 700 by default), over a ripple layer of fixed
depth 0.1 m with sigma 0.072 N/m and rho 1000 kg/m^3.

Against the oracle at M = 1024: the impulse grid and the whitewater (count exactly, records as a bit
for bit multiset: their slots follow atomics) after the droplet stage, the ripple frame to 1e-5, and
whole frames of the coupled FLIP tank with particles bit for bit. The 128^3 workload itself is
checked by properties (determinism, conservation, bounded whitewater)."""
import importlib

import numpy as np
import pytest

import twin
from twin import assert_close, suiri, ww_multiset

workloads = importlib.import_module("hahenn-_amd.workloads")

pytestmark = pytest.mark.gpu

TOL = 1e-5
DT = 1 / 240

LAYER = dict(ripple_depth=0.1, r_sigma_rho=0.072 / 1000.0, r_rho_kick=1000.0, splash_rho=1000.0,
             splash_sigma=0.072, splash_mu=0.001, splash_k_jet=60.0, droplets=4096, drop_seed=2024,
             drop_d50=0.0015, drop_sigma_ln=0.4, drop_vmin=1.0, drop_vmax=4.0)


def per_component(name, got, want, tol):
    for c in range(got.shape[1]):
        if np.abs(want[:, c]).max() > 0:
            assert_close(f"{name}[{c}]", got[:, c], want[:, c], tol)


def layer_twin(R=512, ww_cap=200000, **over):
    extra = dict(LAYER, **over)
    return twin.make_twin(twin.small_scene(R=R, ww_cap=ww_cap, extra=extra))


def test_droplet_stage_and_ripple_frame_match_oracle():
    o, g = layer_twin()
    print("droplets: R=512 M=1024, 4096 per frame, K jet 60", flush=True)
    o.write_params(DT)
    o.L.or_droplets(o.ptr, 0.1)
    g.run_stage(suiri.S_DROPLETS, 0, DT)
    want = o.field("rsrc")
    got = g.readback(suiri.F_RSRC)
    assert np.array_equal(got, want)
    assert want.sum() > 0
    wo = int(o.field("counters")[2])
    gc = g.readback(suiri.F_COUNTERS)
    assert int(gc[2]) == wo and wo > 4000, (gc[2], wo)  # nearly every droplet is above K = 60
    assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo]))
    # the frame that carries the impulses into the ripple state (fixed depth 0.1 m)
    o.L.or_ripple_frame(o.ptr, DT, 0.1)
    g.run_stage(suiri.S_RIPPLE, 0, DT)
    per_component("rstate", g.readback(suiri.F_RSTATE), o.field("rstate"), TOL)
    per_component("rtex", g.readback(suiri.F_RTEX), o.field("tex"), TOL)
    assert np.all(g.readback(suiri.F_RSRC) == 0)


def test_reference_thresholds_spawn_fewer_splashes():
    """With the reference's thresholds (700 / 2100) only the fast or large droplets splash: the same
    draws, fewer whitewater records, the same impulses."""
    o, g = layer_twin(splash_k_jet=700.0, splash_k_crown=2100.0)
    o.write_params(DT)
    o.L.or_droplets(o.ptr, 0.1)
    g.run_stage(suiri.S_DROPLETS, 0, DT)
    wo = int(o.field("counters")[2])
    assert int(g.readback(suiri.F_COUNTERS)[2]) == wo and wo > 0
    assert np.array_equal(g.readback(suiri.F_RSRC), o.field("rsrc"))
    assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo]))


def test_coupled_frames_match_oracle():
    """Whole frames of the FLIP tank with the droplet layer: particles bit for bit, whitewater count
    exact and bit for bit as a multiset (WW moves the splashes of the same frame), ripple state to
    1e-5."""
    o, g = layer_twin()
    for f in range(4):
        print(f"coupled frame {f}", flush=True)
        twin.oracle_step(o, DT)
        twin.gpu_step(g, DT)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
        wo = int(o.info()["ww_count"])
        assert g.stats()["ww"] == wo and wo > 0, f
        assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo])), f
        per_component("rstate", g.readback(suiri.F_RSTATE), o.field("rstate"), TOL)


def test_rgba16f_texture_is_numpy_float16_of_the_f32_texture():
    """The reference's rgba16float ripple texture (64-ripple.js:146, :160): every channel the
    round-to-nearest-even float16 of the f32 texture, as numpy converts it."""
    o, g = layer_twin()
    for _ in range(3):
        twin.gpu_step(g, DT)
    t32 = g.readback(suiri.F_RTEX)
    t16 = g.readback(suiri.F_RTEX16)
    assert np.abs(t32[:, :3]).max() > 0
    assert np.array_equal(t16.view(np.uint16), t32.astype(np.float16).view(np.uint16))


def run_workload(frames):
    wl = workloads.ripple_droplets()
    g = suiri.Sim(suiri.config("high", **wl["cfg"]))
    workloads.setup(g, wl)
    g.reorder()
    g.set_substeps(1)
    n0 = g.stats()["count"]
    for _ in range(frames):
        g.step_items(wl["frame_dt"], wl["items"])
    st = g.stats()
    out = {"n0": n0, "count": st["count"], "ww": st["ww"], "gpu_errors": st["gpuErrors"]}
    out["parts"] = g.readback(suiri.F_PARTICLES)[: st["count"]].view(np.uint32).copy()
    out["rstate"] = g.readback(suiri.F_RSTATE).view(np.uint32).copy()
    out["ww_set"] = ww_multiset(g.readback(suiri.F_WW)[: st["ww"]])
    out["rtex"] = g.readback(suiri.F_RTEX)
    g.close()
    return out


def test_full_size_droplet_workload_properties():
    """The 128^3 tank filled 0.1 m deep with the droplet layer at M = 1024: two runs are bit-identical
    (particles, ripple state, whitewater multiset), the closed tank keeps every particle, whitewater
    accumulates within its capacity, and the ripple height is finite and non-zero."""
    runs = []
    for r in range(2):
        print(f"128^3 droplet workload, run {r}", flush=True)
        runs.append(run_workload(6))
    a, b = runs
    assert a["gpu_errors"] == 0
    assert a["count"] == a["n0"] > 3_000_000
    assert 5 * 4096 <= a["ww"] <= 1_000_000
    for k in ("parts", "rstate", "ww_set"):
        assert np.array_equal(a[k], b[k]), k
    h = a["rtex"][:, 0]
    assert np.all(np.isfinite(a["rtex"])) and np.abs(h).max() > 0
