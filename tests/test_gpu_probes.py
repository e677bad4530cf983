"""Device probes of the library's own device helpers (built by __graft_entry__.build() from
tests/probes/): f2i_hw, the hardware float-to-int conversion G2P's stencil corners use, gives f2i's
truncate-saturate-NaN-to-0 results (WGSL's i32(f), 31-wgsl-sim2.js:192) for every input class."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROBES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probes", "_build")


def test_hardware_float_to_int_matches_f2i():
    exe = os.path.join(PROBES, "cvt_probe")
    assert os.path.exists(exe), "cvt_probe is not built: run __graft_entry__.build()"
    print("cvt_probe: 4 Mi random bit patterns and edge cases", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
