"""Workload generators: seeded, shaped as the stated configurations."""
import numpy as np

from janus_treepm import synth


def test_lattice_zeldovich_shape_and_seed():
    a = synth.lattice_zeldovich(n_side=16, n_cells=16, seed=3)
    b = synth.lattice_zeldovich(n_side=16, n_cells=16, seed=3)
    assert a.n == 2 * 16 ** 3 and np.array_equal(a.pos, b.pos)
    assert np.all(a.pos >= -a.box_size / 2) and np.all(a.pos < a.box_size / 2)
    assert a.softening == a.box_size / a.n ** (1 / 3) / 30
    assert a.r_s == 1.25 * a.box_size / 16 and a.r_cut == 4.5 * a.r_s
    c = synth.lattice_zeldovich(n_side=16, n_cells=16, seed=4)
    assert not np.array_equal(a.pos, c.pos)


def test_clustered_counts():
    w = synth.clustered(n_total=100_000, n_cells=32, n_halos=40, seed=2)
    assert w.n == 100_000 and np.sum(w.species > 0) == 50_000
    assert np.all(np.abs(w.pos) <= w.box_size / 2)


def test_autotune_sub_box_selection():
    """The tuner's sub-box: exactly the particles of a periodic cube of edge L/k, re-centred in [-e/2, e/2)."""
    from janus_treepm.autotune import sub_box
    rng = np.random.default_rng(0)
    L = 100.0
    pos = rng.uniform(-L / 2, L / 2, (20000, 3))
    sp = np.where(rng.random(20000) < 0.5, 1, -1).astype(np.int8)
    x, s, edge, G = sub_box(pos, sp, L, 64, k=4, seed=3)
    assert edge == 25.0 and G == 16
    assert np.all(x >= -edge / 2) and np.all(x < edge / 2)
    assert abs(len(s) - 20000 / 64) < 5 * np.sqrt(20000 / 64)
    lo = np.random.default_rng(3).uniform(-L / 2, L / 2, 3)
    d = (pos - lo) % L
    assert len(s) == int(np.all(d < edge, axis=1).sum())
