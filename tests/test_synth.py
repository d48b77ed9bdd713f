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
