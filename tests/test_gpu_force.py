"""GPU full force: sign rules of the reference tests, edge cases, accuracy against
the FP64 Ewald oracle, and the reference-CPU-compatible mode against the
restatement of TreePMForce (treepm_force.rs)."""
import numpy as np
import pytest

from janus_treepm import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def _rel_rms(a, b):
    """|a - b| over the rms truth magnitude of the sample (no division by near-zero forces)."""
    return np.linalg.norm(a - b, axis=1) / np.sqrt(np.mean(np.sum(b * b, axis=1)))


@pytest.mark.parametrize("si,sj,attract", [(1, 1, True), (-1, -1, True), (1, -1, False), (-1, 1, False)])
@pytest.mark.parametrize("precision", [0, 1])
def test_pair_signs(jtp, si, sj, attract, precision):
    """treepm_force.rs:207-264 / treepm_physics_8p.rs:142-193: (+,+),(-,-) attract, (+,-),(-,+) repel."""
    print(f"case: pair signs {si},{sj} precision={precision}", flush=True)
    f = jtp.TreePMForce(20.0, 32, 100.0, 0.5, 0.5, precision=precision)
    parts = [jtp.Particle((-10.0, 0.0, 0.0), 1.0, si), jtp.Particle((10.0, 0.0, 0.0), 1.0, sj)]
    f.update(parts)
    a = f.compute_all_forces(parts)
    assert (a[0, 0] > 0) == attract
    assert (a[1, 0] < 0) == attract
    assert abs(a[0, 0] + a[1, 0]) < 0.01 * abs(a[0, 0])  # Newton's third law (equal masses)


def test_edge_cases(jtp):
    print("case: edge cases", flush=True)
    p = jtp.default_params(100.0, 32)
    ctx = jtp.ForceContext(p)
    # n = 0 is a no-op
    ctx.compute_forces(np.zeros((0, 3), np.uint32), np.zeros(0, np.int8))
    # n = 1: no short-range force, PM force of a lone particle vanishes by symmetry
    a = ctx.compute_forces(np.array([[123, 456, 789]], np.uint32), np.array([1], np.int8))
    assert np.all(np.abs(a) < 1e-6)
    # coincident particles: finite (softened) and zero by symmetry
    pos = np.tile(np.array([[2 ** 31, 2 ** 31, 2 ** 31]], np.uint32), (5, 1))
    a = ctx.compute_forces(pos, np.array([1, 1, -1, -1, 1], np.int8))
    assert np.all(np.isfinite(a)) and np.all(np.abs(a) < 1e-4)
    # particles on the periodic boundary (u = 0 and u = 2^32 - 1 are neighbours)
    pos = np.array([[0, 0, 0], [2 ** 32 - 1, 0, 0]], np.uint32)
    a = ctx.compute_forces(pos, np.array([1, 1], np.int8))
    assert np.all(np.isfinite(a))
    # invalid arguments are rejected with JTP_ERR_ARG, never a crash
    lib = jtp.load_library()
    assert lib.jtp_compute_forces(ctx.h, -1, None, None, None, None) == jtp.JTP_ERR_ARG
    assert lib.jtp_compute_forces(ctx.h, 10, None, None, None, None) == jtp.JTP_ERR_ARG
    bad = jtp.default_params(100.0, 48)
    with pytest.raises(jtp.JtpError):
        jtp.ForceContext(bad)


@pytest.mark.parametrize("precision", [0, 1])
def test_full_force_vs_ewald(jtp, orc, precision):
    print(f"case: full force vs Ewald precision={precision}", flush=True)
    w = synth.lattice_zeldovich(n_side=24, n_cells=32, seed=21)
    pos = jtp.to_fixed(w.pos, w.box_size)
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.theta, p.softening, p.precision = w.r_s, w.r_cut, 0.5, w.softening, precision
    ctx = jtp.ForceContext(p, w.m_plus, w.m_minus_abs)
    acc = ctx.compute_forces(pos, w.species).astype(np.float64)
    rng = np.random.default_rng(0)
    tg = rng.choice(w.n, 400, replace=False)
    ref = orc.ewald(pos, w.species, w.m_plus, w.m_minus_abs, w.box_size, w.softening, targets=tg, rc=w.box_size / 3)
    e = _rel(acc[tg], ref)
    er = _rel_rms(acc[tg], ref)
    print("TreePM vs Ewald (2x24^3, G=32): median %.2e p99 %.2e; |da| / rms|a|: median %.2e p99 %.2e"
          % (np.median(e), np.percentile(e, 99), np.median(er), np.percentile(er, 99)))
    # a regression bound on a small near-uniform lattice (measured: 6.9e-3 / 4.8e-2 relative, 4.3e-3 /
    # 1.7e-2 of the rms force); the accuracy bar is asserted on the clustered boxes in test_gpu_headline.py
    assert np.median(e) < 1e-2 and np.percentile(e, 99) < 6e-2
    assert np.median(er) < 6e-3 and np.percentile(er, 99) < 2.5e-2


def test_compat_vs_reference_cpu(jtp, orc):
    """Reference-CPU-compatible switches (x^4 split, 2-point gradient, no deconvolution, r_s = r_cut/3)
    at theta = 0 against the restatement of TreePMForce; particles sit in the inner half of the box so the
    reference's non-periodic tree sees the same pairs as the minimum-image walk."""
    print("case: compat vs reference CPU", flush=True)
    from oracle.oracle_lib import RefCpu
    L, G = 100.0, 32
    rng = np.random.default_rng(5)
    n = 3000
    x = rng.uniform(-L / 4, L / 4, (n, 3))
    sp = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int8)
    pos = jtp.to_fixed(x, L)
    xq = jtp.from_fixed(pos, L)
    r_cut, eps = 8.0, 0.3
    p = jtp.default_params(L, G)
    p.r_cut, p.r_s, p.theta, p.softening = r_cut, r_cut / 3, 0.0, eps
    p.precision, p.split_mode, p.gradient, p.deconvolve, p.interlace = 1, 2, 1, 0, 0
    ctx = jtp.ForceContext(p, 1.0, 1.0)
    acc = ctx.compute_forces(pos, sp)
    h = L / G
    ref = RefCpu(orc, r_cut, G, L, 0.0, eps, r_cut / 3, 1.0 / h ** 3, 1.0)
    ref.update(xq, np.ones(n), sp)
    a_ref = ref.forces()
    e = _rel(acc, a_ref)
    print("compat vs reference CPU: median %.2e max %.2e" % (np.median(e), np.max(e)))
    assert np.median(e) < 1e-9 and np.max(e) < 1e-6


@pytest.mark.parametrize("precision", [1, 0])
def test_config0_full_force_and_step(jtp, orc, precision):
    """The small configuration: N = 2 x 64^3 lattice + Zel'dovich (P(k) ~ k^-2.5, rms 0.3 cell, m- on the
    half-cell-offset lattice with the opposite displacement), L = 250, PM 64^3, r_s = 1.25 cells,
    r_cut = 4.5 r_s, theta = 0.5, leaf 8, m+ = 1, |m-| = 19: one force evaluation against the FP64 Ewald
    sum on 1000 targets, then one DKD step (finite, momentum of the kick conserved).

    A near-uniform lattice's forces are small residuals of cancelling contributions, and at r_s = 1.25
    cells its error is set by the mesh part: theta = 0 (no tree approximation) still gives a relative
    median of 3.6e-3, theta = 0.5 4.1e-3, interlacing off 1.8e-2 (measured).  So the per-particle
    relative bar of the headline box (1e-3 / 1e-2, clustered, test_gpu_headline.py) is not this
    configuration's; it is held to the error over the rms force of the sample, |da| / rms|a|: median
    <= 2.5e-3 and 99th percentile <= 1e-2 (measured 1.6e-3 / 4.1e-3), and to a relative median
    <= 6e-3, p99 <= 5e-2 (measured 4.1e-3 / 2.3e-2)."""
    print(f"case: config 0 precision={precision}", flush=True)
    w = synth.lattice_zeldovich(n_side=64, box_size=250.0, n_cells=64, seed=42, leaf_size=8)
    assert w.n == 2 * 64 ** 3
    pos = jtp.to_fixed(w.pos, w.box_size)
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.softening, p.leaf_size, p.precision = w.r_s, w.r_cut, w.softening, 8, precision
    ctx = jtp.ForceContext(p, w.m_plus, w.m_minus_abs)
    acc = ctx.compute_forces(pos, w.species).astype(np.float64)
    rng = np.random.default_rng(7)
    tg = rng.choice(w.n, 1000, replace=False)
    ref = orc.ewald(pos, w.species, w.m_plus, w.m_minus_abs, w.box_size, w.softening, targets=tg,
                    rc=w.box_size / 3)
    e = _rel(acc[tg], ref)
    er = _rel_rms(acc[tg], ref)
    print("config 0 vs Ewald: median %.2e p99 %.2e; |da| / rms|a|: median %.2e p99 %.2e"
          % (np.median(e), np.percentile(e, 99), np.median(er), np.percentile(er, 99)))
    assert np.median(er) <= 2.5e-3 and np.percentile(er, 99) <= 1e-2
    assert np.median(e) <= 6e-3 and np.percentile(e, 99) <= 5e-2
    rt = np.float64 if precision == 1 else np.float32
    n = w.n
    dp, dv = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(np.zeros((n, 3), rt))
    ds, da = jtp.DeviceArray.from_host(w.species), jtp.DeviceArray((n, 3), rt)
    ctx.step_device(n, dp, dv, ds, da, 0.01, scheme=jtp.STEP_DKD)
    v = dv.download().astype(np.float64)
    assert np.all(np.isfinite(v))
    m = np.where(w.species > 0, w.m_plus, w.m_minus_abs)[:, None]
    mom = (m * v).sum(axis=0)
    assert np.abs(mom).max() < 1e-4 * (m * np.abs(v)).sum()
