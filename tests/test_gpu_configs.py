"""The two large configurations at full size on the GPU.

Ensemble: the 48 lattice + Zel'dovich boxes of synth.ensemble_c3 (N_b = 2 n^3 log2-uniform in
[2 x 32^3, 2 x 128^3], 3.98e7 particles in total, PM 32^3 - 128^3, L ~ U[100, 500], |m-|/m+ ~ U[5, 40]):
every box's FP32 accelerations from jtp_ensemble_compute_forces are bit-identical to a single-context
jtp_compute_forces of that box, in every ensemble mode (one walk launch over all boxes or one per box).

Large box: N = 2 x 256^3 + 7.2e6 m+ = 40,754,432 particles of the clustered recipe with 8000 halos,
L = 300, PM 512^3, theta = 0.5: per-species deposited mass conservation, bit-identical reruns, the FP32
form against the plain FP64 form, and three graph-replayed KDK steps (FP64 drift) whose momentum stays
conserved.  Reference step: step_treepm_gpu, nbody_gpu_twopass.rs:4486-4649."""
import numpy as np
import pytest

from janus_treepm import synth

pytestmark = pytest.mark.gpu


def _params(jtp, w, precision=0, **kw):
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.theta, p.softening, p.leaf_size = w.r_s, w.r_cut, w.theta, w.softening, w.leaf_size
    p.precision = precision
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_ensemble_48_boxes_match_single_boxes(jtp):
    print("case: 48-box ensemble vs single-box contexts", flush=True)
    ws = synth.ensemble_c3()
    assert len(ws) == 48
    n = [w.n for w in ws]
    assert sum(n) > 3.9e7 and min(n) >= 2 * 32 ** 3 and max(n) <= 2 * 128 ** 3
    dpos = [jtp.DeviceArray.from_host(jtp.to_fixed(w.pos, w.box_size)) for w in ws]
    dsp = [jtp.DeviceArray.from_host(w.species) for w in ws]
    dacc = [jtp.DeviceArray((w.n, 3), np.float32) for w in ws]
    ref = []
    for b, w in enumerate(ws):
        ctx = jtp.ForceContext(_params(jtp, w), w.m_plus, w.m_minus_abs)
        ctx.compute_forces_device(n[b], dpos[b], dsp[b], dacc[b])
        ref.append(dacc[b].download())
        ctx.close()
        assert np.all(np.isfinite(ref[-1]))
    ens = jtp.Ensemble([_params(jtp, w) for w in ws], [w.m_plus for w in ws], [w.m_minus_abs for w in ws],
                       n_streams=4)
    for mode in ("concurrent", "sequential", "graph", "graph", "per_box", "per_box graph", "per_box graph"):
        print(f"  mode {mode}", flush=True)
        for d in dacc:
            d.upload(np.full((d.shape[0], 3), np.nan, np.float32))
        ens.compute_forces_device(n, dpos, dsp, dacc, sequential="sequential" in mode, graph="graph" in mode,
                                  per_box_walk="per_box" in mode)
        for b, d in enumerate(dacc):
            assert np.array_equal(d.download(), ref[b]), f"mode {mode} box {b}"
    ens.close()


@pytest.fixture(scope="module")
def box41m():
    n_plus = 256 ** 3 + 7_200_000
    w = synth.clustered(n_total=n_plus + 256 ** 3, n_plus=n_plus, box_size=300.0, n_cells=512, n_halos=8000,
                        seed=11)
    assert w.n == 40_754_432
    return w


def test_large_box_mass_and_rerun(jtp, box41m):
    """Per-species mass on the 512^3 grids (m- mass set to 0, then m+ mass set to 0: the net grid sums to
    the other species' total), and two evaluations giving the same bits, for both forms."""
    w = box41m
    print(f"case: 41M box mass conservation and reruns, N={w.n}", flush=True)
    dpos, dsp = jtp.DeviceArray.from_host(jtp.to_fixed(w.pos, w.box_size)), jtp.DeviceArray.from_host(w.species)
    n_plus, n_minus = int(np.sum(w.species > 0)), int(np.sum(w.species < 0))
    for precision in (0, 1):
        rt = np.float64 if precision == 1 else np.float32
        tol = 1e-12 if precision == 1 else 1e-5
        ctx = jtp.ForceContext(_params(jtp, w, precision, tree_enabled=0), w.m_plus, w.m_minus_abs)
        for mp, mm in ((w.m_plus, 0.0), (0.0, w.m_minus_abs)):
            ctx.set_species_mass(mp, mm)
            da = jtp.DeviceArray((w.n, 3), rt)
            ctx.compute_forces_device(w.n, dpos, dsp, da)
            for st in (jtp.STAGE_RHO, jtp.STAGE_RHO2):
                tot = float(np.sum(ctx.stage(st).astype(np.float64)))
                want = mp * n_plus - mm * n_minus
                print(f"  precision {precision} masses ({mp}, {mm}) stage {st}: sum {tot:.9e} want {want:.9e}",
                      flush=True)
                assert abs(tot - want) <= tol * abs(want)
            da.free()
        ctx.close()
        ctx = jtp.ForceContext(_params(jtp, w, precision), w.m_plus, w.m_minus_abs)
        a1, a2 = jtp.DeviceArray((w.n, 3), rt), jtp.DeviceArray((w.n, 3), rt)
        ctx.compute_forces_device(w.n, dpos, dsp, a1)
        ctx.compute_forces_device(w.n, dpos, dsp, a2)
        h1 = a1.download()
        assert np.all(np.isfinite(h1)) and np.array_equal(h1, a2.download())
        ctx.close()
        a1.free()
        a2.free()


def test_large_box_forms_and_kdk_steps(jtp, box41m):
    """FP32 form against the FP64 form at identical parameters, then three graph-replayed KDK steps of the
    FP32 form from rest: the total momentum sum_s sum |m| v (conserved by the Janus pair law) stays below
    1e-3 of sum |m| |v|."""
    w = box41m
    print(f"case: 41M box FP32 vs FP64 and KDK steps, N={w.n}", flush=True)
    pos = jtp.to_fixed(w.pos, w.box_size)
    dpos, dsp = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(w.species)
    acc = {}
    for precision in (1, 0):
        rt = np.float64 if precision == 1 else np.float32
        ctx = jtp.ForceContext(_params(jtp, w, precision), w.m_plus, w.m_minus_abs)
        da = jtp.DeviceArray((w.n, 3), rt)
        ctx.compute_forces_device(w.n, dpos, dsp, da)
        acc[precision] = da.download().astype(np.float64)
        ctx.close()
        da.free()
    e = np.linalg.norm(acc[0] - acc[1], axis=1) / np.maximum(np.linalg.norm(acc[1], axis=1), 1e-300)
    print("41M FP32 vs FP64: median %.2e p99 %.2e" % (np.median(e), np.percentile(e, 99)), flush=True)
    assert np.median(e) < 1e-4 and np.percentile(e, 99) < 1e-2
    a1 = acc[1]
    del acc
    # KDK steps with dt such that the p99 particle moves ~0.1 cell per step
    h = w.box_size / w.n_cells
    dt = float(np.sqrt(2.0 * 0.1 * h / np.percentile(np.linalg.norm(a1, axis=1), 99)))
    del a1
    ctx = jtp.ForceContext(_params(jtp, w, 0), w.m_plus, w.m_minus_abs)
    dv = jtp.DeviceArray.from_host(np.zeros((w.n, 3), np.float32))
    da = jtp.DeviceArray((w.n, 3), np.float32)
    ctx.compute_forces_device(w.n, dpos, dsp, da)
    for k in range(3):
        ctx.step_device(w.n, dpos, dv, dsp, da, dt, scheme=jtp.STEP_KDK, graph=True)
        print(f"  KDK step {k + 1} done", flush=True)
    v = dv.download().astype(np.float64)
    assert np.all(np.isfinite(v))
    m = np.where(w.species > 0, w.m_plus, w.m_minus_abs)[:, None]
    mom = (m * v).sum(axis=0)
    scale = (m * np.abs(v)).sum()
    print("momentum drift / sum |p|: %.2e" % (np.abs(mom).max() / scale), flush=True)
    assert np.abs(mom).max() < 1e-3 * scale
    d = ctx.diagnostics(w.n, dv, dsp)
    assert np.allclose(d["momentum"].sum(axis=0), mom, rtol=1e-6, atol=1e-9 * scale)
    ctx.close()
