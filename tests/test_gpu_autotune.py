"""Autotuner of leaf / group size and FFT tile: runs on a sub-box, returns a candidate no less
accurate (within the tolerance) than the configured one and no slower than it, and the faster FFT
tile; every FFT tile gives bit-identical results."""
import numpy as np
import pytest

from janus_treepm import synth

pytestmark = pytest.mark.gpu


def test_autotune_sub_box(jtp):
    print("case: autotune", flush=True)
    from janus_treepm.autotune import autotune
    w = synth.clustered(n_total=400_000, box_size=100.0, n_cells=64, n_halos=60, seed=41)
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.softening = w.r_s, w.r_cut, w.softening
    res = autotune(w.pos, w.species, p, w.m_plus, w.m_minus_abs, k=2,
                   candidates=((4, 64), (8, 64), (4, 32)), repeats=2)
    print(res)
    assert res["sample"] > 10_000
    rows = {(r[0], r[1]): r for r in res["table"]}
    base = rows[(p.leaf_size, p.group_size)]
    pick = rows[(res["leaf_size"], res["group_size"])]
    assert pick[3] <= 1.05 * base[3]
    assert pick[2] <= base[2]
    assert all(np.isfinite(r[2]) and 0 < r[3] < 1e-2 for r in res["table"])
    assert res["fft_tile"] in (11, 12, 13) and len(res["fft_table"]) == 3


@pytest.mark.parametrize("G", [32, 128, 256])
def test_fft_tiles_bit_identical(jtp, G):
    print(f"case: FFT tiles 11, 12, 13 G={G}", flush=True)
    rng = np.random.default_rng(G)
    rho = rng.standard_normal((G, G, G)).astype(np.float32)
    out = []
    for tile in (11, 12, 13):
        p = jtp.default_params(100.0, G)
        p.r_cut = 100.0 / 8
        p.fft_tile = tile
        ctx = jtp.ForceContext(p)
        d_rho, d_phi = jtp.DeviceArray.from_host(rho), jtp.DeviceArray(rho.shape, np.float32)
        ctx.pm_solve_device(d_rho, d_phi)
        out.append(d_phi.download())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    w = synth.clustered(n_total=200_000, box_size=100.0, n_cells=64, n_halos=30, seed=5)
    acc = []
    for tile in (11, 13):
        p = jtp.default_params(w.box_size, w.n_cells)
        p.fft_tile = tile
        acc.append(jtp.ForceContext(p, w.m_plus, w.m_minus_abs).compute_forces(jtp.to_fixed(w.pos, w.box_size),
                                                                              w.species))
    assert np.array_equal(acc[0], acc[1])
