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


def test_default_params_headline_split():
    """jtp_default_params: the split of the headline workloads (r_s = 1.25 cells, r_cut = 4.5 r_s), not the
    reference's derivation (that is jtp_params_from_config without split fields, tested below)."""
    p = j.default_params(250.0, 64)
    cell = 250.0 / 64
    assert p.r_s == pytest.approx(1.25 * cell) and p.r_cut == pytest.approx(4.5 * 1.25 * cell)
    assert p.theta == 0.5 and p.leaf_size == 4 and p.group_size == 32 and p.precision == j.FP32
    assert p.interlace == 1 and p.theta_cut == pytest.approx(0.33) and p.deconvolve == 1
    assert p.fft_tile == 0


def test_config_fields_follow_reference_derivation():
    """Without the optional split fields the YAML helper reproduces the reference: r_cut = 2 L / n_cells
    (janus_optim.rs:74), r_s = r_cut / 3 (nbody_gpu_twopass.rs:1694), theta defaults to 0.7 (config.rs:42),
    softening to box / N^(1/3) / 30 (config.rs:191-196), k_min passes through."""
    L, G, N = 500.0, 128, 2_000_000
    p = j.params_from_config(L, G, n_particles=N, k_min=2)
    assert p.r_cut == 2.0 * L / G
    assert p.r_s == pytest.approx(p.r_cut / 3.0, rel=1e-15)
    assert p.theta == 0.7 and p.k_min == 2 and p.n_cells == G and p.box_size == L
    assert p.softening == L / N ** (1.0 / 3.0) / 30.0
    # the rest as jtp_default_params
    d = j.default_params(L, G)
    for f in ("leaf_size", "group_size", "precision", "split_mode", "gradient", "deconvolve", "interlace"):
        assert getattr(p, f) == getattr(d, f), f
    # explicit fields win; the headline split comes back from split_scale_cells = 1.25, r_cut_over_r_s = 4.5
    q = j.params_from_config(250.0, 256, n_particles=22_000_000, theta=0.5, softening_mpc=0.03,
                             split_scale_cells=1.25, r_cut_over_r_s=4.5)
    h = j.default_params(250.0, 256)
    assert q.r_s == h.r_s and q.r_cut == pytest.approx(h.r_cut, rel=1e-15)
    assert q.theta == 0.5 and q.softening == 0.03 and q.theta_cut == pytest.approx(0.33)
    # only the ratio given: r_cut stays at 2 cells, r_s = r_cut / ratio
    r = j.params_from_config(L, G, n_particles=N, r_cut_over_r_s=4.0)
    assert r.r_cut == pytest.approx(2.0 * L / G, rel=1e-15) and r.r_s == pytest.approx(0.5 * L / G, rel=1e-15)
    # invalid: no softening and no particle count; a split beyond L / 5
    with pytest.raises(j.JtpError):
        j.params_from_config(L, G)
    with pytest.raises(j.JtpError):
        j.params_from_config(L, 16, n_particles=N, split_scale_cells=2.0, r_cut_over_r_s=4.5)
    lib = j.load_library()
    assert lib.jtp_params_from_config(None, None) == j.JTP_ERR_ARG


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
    assert lib.jtp_morton30_keys_f32(-1, None, 0.0, 0.0, None, None, None) == j.JTP_ERR_ARG
    assert lib.jtp_morton30_keys_f32(2 ** 31, None, 0.0, 0.0, None, None, None) == j.JTP_ERR_ARG
    assert lib.jtp_morton30_keys_f32(10, None, 0.0, 0.0, None, None, None) == j.JTP_ERR_ARG
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


def test_device_array_arguments_checked():
    """compute_forces_device / step_device refuse a DeviceArray of the wrong dtype or size before the
    library is called (a float32 acceleration buffer on an FP64 context would be written past its end)."""
    ctx = j.ForceContext.__new__(j.ForceContext)
    ctx.lib, ctx.h = j.load_library(), None
    ctx.params = j.default_params(100.0, 32)
    ctx.params.precision = j.FP64

    def fake(shape, dtype):
        d = j.DeviceArray.__new__(j.DeviceArray)
        d.shape, d.dtype = shape, np.dtype(dtype)
        d.nbytes = int(np.prod(shape)) * d.dtype.itemsize
        d.ptr = None  # never dereferenced: the check comes first
        return d

    assert ctx._arg(fake((10, 3), np.float64), ctx.real, 30, "acc") is None
    assert ctx._arg(12345, ctx.real, 30, "acc") == 12345  # raw pointers pass through unchecked
    with pytest.raises(ValueError):
        ctx._arg(fake((10, 3), np.float32), ctx.real, 30, "acc")
    with pytest.raises(ValueError):
        ctx._arg(fake((9, 3), np.float64), ctx.real, 30, "acc")
    with pytest.raises(ValueError):
        ctx._arg(fake((10, 3), np.int32), np.uint32, 30, "pos")
    with pytest.raises(ValueError):
        ctx.compute_forces_device(10, fake((10, 3), np.uint32), fake((10,), np.int8), fake((10, 3), np.float32))
    with pytest.raises(ValueError):
        ctx.step_device(10, fake((10, 3), np.uint32), fake((10, 3), np.float64), fake((5,), np.int8),
                        fake((10, 3), np.float64), 0.01)
