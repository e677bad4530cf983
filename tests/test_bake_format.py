"""Bake map format (the layout of the reference's exr.py:11-41), read back with an OpenEXR reader
written from the file-format specification."""
import importlib

import numpy as np

import exrspec

suiri = importlib.import_module("hahenn-_amd")


def test_exr_maps_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    H, W = 17, 23
    r, g, b = (rng.standard_normal((H, W)).astype(np.float32) for _ in range(3))
    r[0, 0] = -1.0   # negative values are kept exactly (dry texels)
    path = tmp_path / "surface_00003.exr"
    suiri.write_exr_rgb(str(path), r, g, b)
    attrs, chans, planes = exrspec.read(str(path))
    assert chans == ["B", "G", "R"]
    assert attrs["lineOrder"][1] == b"\0"
    assert np.array_equal(planes["R"], r) and np.array_equal(planes["G"], g) and np.array_equal(planes["B"], b)
