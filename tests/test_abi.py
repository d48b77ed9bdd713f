"""CPU checks of the C ABI: the library loads without a GPU, exports every
function declared in include/janus_treepm.h, and rejects invalid parameters
with JTP_ERR_ARG before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import janus_treepm as j

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "janus_treepm.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jtp_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = j.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(j.EXPORTED) == names


def test_default_params_follow_reference_derivation():
    p = j.default_params(250.0, 64)
    cell = 250.0 / 64
    assert p.r_s == pytest.approx(1.25 * cell) and p.r_cut == pytest.approx(4.5 * 1.25 * cell)
    assert p.theta == 0.5 and p.leaf_size == 4 and p.group_size == 32 and p.precision == j.FP32
    assert p.interlace == 1 and p.theta_cut == pytest.approx(0.33) and p.deconvolve == 1
    assert p.fft_tile == 0


@pytest.mark.parametrize("field,value", [
    ("n_cells", 48), ("n_cells", 8), ("n_cells", 1024), ("box_size", 0.0), ("r_cut", 60.0), ("r_cut", 0.0),
    ("theta", 1.5), ("softening", 0.0), ("leaf_size", 0), ("leaf_size", 17), ("group_size", 65),
    ("precision", 2), ("split_mode", 3), ("gradient", 2), ("k_min", -1), ("r_s", 0.0), ("theta_cut", -0.1), ("interlace", 2),
    ("fft_tile", 10), ("fft_tile", 14),
])
def test_create_rejects_bad_params(field, value):
    lib = j.load_library()
    p = j.default_params(250.0, 64)
    setattr(p, field, value)
    h = C.c_void_p()
    assert lib.jtp_create(C.byref(p), C.byref(h)) == j.JTP_ERR_ARG
    assert not h.value


def test_null_arguments():
    lib = j.load_library()
    assert lib.jtp_create(None, None) == j.JTP_ERR_ARG
    assert lib.jtp_compute_forces(None, 0, None, None, None, None) == j.JTP_ERR_ARG
    assert lib.jtp_set_species_mass(None, 1.0, 1.0) == j.JTP_ERR_ARG
    assert lib.jtp_morton_keys_f32(-1, None, 0.0, 0.0, None, None) == j.JTP_ERR_ARG
    assert lib.jtp_destroy(None) == j.JTP_OK


def test_fixed_point_round_trip():
    rng = np.random.default_rng(0)
    L = 250.0
    x = rng.uniform(-L / 2, L / 2, (1000, 3))
    u = j.to_fixed(x, L)
    back = j.from_fixed(u, L)
    assert np.max(np.abs(back - x)) <= L / 2 ** 32
    assert np.array_equal(j.to_fixed(x + L, L), u)
    assert np.array_equal(j.to_fixed(np.array([-L / 2, 0.0, L / 2 - 1e-12]), L)[:2], [0, 2 ** 31])


def test_fixed_point_matches_oracle(orc):
    rng = np.random.default_rng(1)
    x = rng.uniform(-400, 400, (5000, 3))
    assert np.array_equal(j.to_fixed(x, 250.0), orc.fixed(x, 250.0))
