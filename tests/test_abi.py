"""The C-ABI library loads, exports every function include/suiri.h declares, fills the reference
defaults, and reports errors without a device (no compute is launched here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "suiri.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sl_[a-z_0-9]+)\s*\(", txt)) - {"sl_emit_fn"})


def test_library_exports_every_declared_function(suiri):
    L = suiri.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert sorted(suiri.EXPORTS) == names


def test_config_defaults_are_reference_quality(suiri):
    # QUALITY, src/js/50-sim.js:4-8 and writeParams constants, :157-160
    c = suiri.config("low")
    assert (c.h, c.cap, c.ww_cap, c.iters, c.ripple_R) == (0.012, 140000, 60000, 28, 128)
    c = suiri.config("high")
    assert (c.h, c.cap, c.ww_cap, c.iters, c.ripple_R) == (0.0065, 700000, 260000, 56, 512)
    c = suiri.config("standard")
    assert (c.h, c.cap, c.ww_cap, c.iters, c.ripple_R) == (0.0085, 320000, 140000, 40, 256)
    assert c.flip == pytest.approx(0.96) and c.omega == pytest.approx(1.6) and c.rc_factor == 0.32
    assert c.ppc == 8 and c.drift == 6 and c.friction == pytest.approx(0.03) and c.ww_rate == 7
    with pytest.raises(suiri.SuiriError):
        suiri.config("ultra")


def test_errors_are_reported(suiri):
    L = suiri.load()
    c = suiri.config("standard", h=-1.0)
    p = ctypes.c_void_p()
    assert L.sl_create(ctypes.byref(c), ctypes.byref(p)) != 0
    assert not p.value
    assert len(L.sl_last_error(None)) > 0
    # null arguments are rejected before any device work
    assert L.sl_stats(None, None) == -1
    assert L.sl_sync(None) == -1
