"""GPU ensembles (jtp_ensemble_*): every box's forces equal jtp_compute_forces on that box alone,
bit for bit, whether the boxes are prepared concurrently over internal streams or sequentially, walked
by one launch over all boxes' groups (the default; also with boxes of different group sizes, three
launches, and a box without a tree) or box by box (JTP_ENS_WALK_PER_BOX), eagerly or as a replayed
graph of the whole batch; empty boxes and uneven parameters included."""
import numpy as np
import pytest

from janus_treepm import synth

pytestmark = pytest.mark.gpu


def _boxes():
    specs = [(12, 100.0, 32, 7.0), (20, 300.0, 32, 19.0), (16, 180.0, 64, 30.0), (24, 450.0, 64, 5.0),
             (10, 120.0, 32, 12.0)]
    ws = []
    for b, (ns, L, G, ratio) in enumerate(specs):
        ws.append(synth.lattice_zeldovich(n_side=ns, box_size=L, n_cells=G, seed=500 + b, m_minus_abs=ratio))
    return ws


def _params(jtp, w, precision):
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.softening, p.precision = w.r_s, w.r_cut, w.softening, precision
    return p


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_matches_single_boxes(jtp, precision):
    print(f"case: ensemble vs single boxes precision={precision}", flush=True)
    ws = _boxes()
    rt = np.float64 if precision == 1 else np.float32
    ref = []
    for w in ws:
        ctx = jtp.ForceContext(_params(jtp, w, precision), w.m_plus, w.m_minus_abs)
        ref.append(ctx.compute_forces(jtp.to_fixed(w.pos, w.box_size), w.species))
        ctx.close()
    ens = jtp.Ensemble([_params(jtp, w, precision) for w in ws], [w.m_plus for w in ws],
                       [w.m_minus_abs for w in ws], n_streams=3)
    dpos = [jtp.DeviceArray.from_host(jtp.to_fixed(w.pos, w.box_size)) for w in ws]
    dsp = [jtp.DeviceArray.from_host(w.species) for w in ws]
    dacc = [jtp.DeviceArray((w.n, 3), rt) for w in ws]
    n = [w.n for w in ws]
    for mode in ["concurrent", "sequential", "graph", "graph", "per_box", "per_box sequential",
                 "per_box graph", "per_box graph", "concurrent", "graph"]:
        for d in dacc:
            d.upload(np.full((d.shape[0], 3), np.nan, rt))
        ens.compute_forces_device(n, dpos, dsp, dacc, sequential="sequential" in mode, graph="graph" in mode,
                                  per_box_walk="per_box" in mode)
        for b, d in enumerate(dacc):
            assert np.array_equal(d.download(), ref[b]), f"mode {mode} box {b}"
        st = ens.stats()
        assert st["n_particles"] == sum(n) and st["ms_total"] > 0 and st["interactions"] > 0
    # an empty box is skipped; the others are unchanged
    n2 = list(n)
    n2[1] = 0
    ens.compute_forces_device(n2, dpos, dsp, dacc, graph=True)
    for b in (0, 2, 3, 4):
        assert np.array_equal(dacc[b].download(), ref[b])
    ens.compute_forces_device(n2, dpos, dsp, dacc, per_box_walk=True)
    for b in (0, 2, 3, 4):
        assert np.array_equal(dacc[b].download(), ref[b])
    ens.close()


def test_one_walk_mixed_boxes(jtp):
    """Boxes with groups of 16, 32 and 64 (three walk launches) and a box without a tree (PM only)."""
    print("case: one-walk ensemble with mixed group sizes and a PM-only box", flush=True)
    ws = _boxes()
    params = [_params(jtp, w, 0) for w in ws]
    params[0].group_size, params[1].group_size, params[3].group_size = 16, 64, 16
    params[2].tree_enabled = 0
    ref = []
    for w, p in zip(ws, params):
        ctx = jtp.ForceContext(p, w.m_plus, w.m_minus_abs)
        ref.append(ctx.compute_forces(jtp.to_fixed(w.pos, w.box_size), w.species))
        ctx.close()
    ens = jtp.Ensemble(params, [w.m_plus for w in ws], [w.m_minus_abs for w in ws], n_streams=2)
    dpos = [jtp.DeviceArray.from_host(jtp.to_fixed(w.pos, w.box_size)) for w in ws]
    dsp = [jtp.DeviceArray.from_host(w.species) for w in ws]
    dacc = [jtp.DeviceArray((w.n, 3), np.float32) for w in ws]
    n = [w.n for w in ws]
    for graph in (False, True, True):
        for d in dacc:
            d.upload(np.full((d.shape[0], 3), np.nan, np.float32))
        ens.compute_forces_device(n, dpos, dsp, dacc, graph=graph)
        for b, d in enumerate(dacc):
            assert np.array_equal(d.download(), ref[b]), f"graph {graph} box {b}"
        st = ens.stats()
        assert st["interactions"] > 0
        for b in (0, 1, 3, 4):
            assert ens.stats(b)["interactions"] > 0
    ens.close()


def test_ensemble_rejects_bad_arguments(jtp):
    print("case: ensemble argument checks", flush=True)
    ws = _boxes()[:2]
    ens = jtp.Ensemble([_params(jtp, w, 0) for w in ws], [1.0, 1.0], [1.0, 1.0])
    with pytest.raises(jtp.JtpError):
        ens.compute_forces_device([-1, 5], [0, 0], [0, 0], [0, 0])
    with pytest.raises(jtp.JtpError):
        ens.stats(5)
    # device arrays of the wrong type or size are refused before any kernel runs
    dpos = [jtp.DeviceArray.from_host(jtp.to_fixed(w.pos, w.box_size)) for w in ws]
    dsp = [jtp.DeviceArray.from_host(w.species) for w in ws]
    n = [w.n for w in ws]
    wrong_acc = [jtp.DeviceArray((w.n, 3), np.float64) for w in ws]  # FP32 boxes
    short_acc = [jtp.DeviceArray((w.n - 1, 3), np.float32) for w in ws]
    with pytest.raises(ValueError):
        ens.compute_forces_device(n, dpos, dsp, wrong_acc)
    with pytest.raises(ValueError):
        ens.compute_forces_device(n, dpos, dsp, short_acc)
    with pytest.raises(ValueError):
        ens.compute_forces_device(n[:1], dpos, dsp, short_acc)
    for a in dpos + dsp + wrong_acc + short_acc:
        a.free()
    ens.close()
    bad = _params(jtp, ws[0], 0)
    bad.n_cells = 48
    with pytest.raises(jtp.JtpError):
        jtp.Ensemble([bad], [1.0], [1.0])
