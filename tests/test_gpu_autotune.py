"""Autotuner of leaf / group size: runs on a sub-box, returns a candidate no less accurate (within the
tolerance) than the configured one and no slower than it."""
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
