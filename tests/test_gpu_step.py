"""GPU integrator step (jtp_step): the reference's step_treepm_gpu order (drift dt/2, force, kick dt with
Hubble friction, drift dt/2; nbody_gpu_twopass.rs:4486-4649 with kick_f32 / drift_f32 :33-64), the KDK
variant, graph replay against eager launches, closed forms for the friction kick, the reference's
two-body energy test (treepm_physics_8p.rs:267-374), momentum conservation and time reversibility."""
import numpy as np
import pytest

from janus_treepm import synth

pytestmark = pytest.mark.gpu

TWO32 = 4294967296.0


def _ctx(jtp, w, precision, **kw):
    p = jtp.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.softening, p.precision = w.r_s, w.r_cut, w.softening, precision
    for k, v in kw.items():
        setattr(p, k, v)
    return jtp.ForceContext(p, w.m_plus, w.m_minus_abs)


def _host_drift(u, v, dt, L):
    """drift on the host with the device's integer arithmetic: u += rint(v dt 2^32 / L) mod 2^32."""
    d = v.astype(np.float64) * (dt * (TWO32 / L))
    d -= TWO32 * np.rint(d / TWO32)
    return ((u.astype(np.int64) + np.rint(d).astype(np.int64)) % (1 << 32)).astype(np.uint32)


def _state(jtp, w, rng, vscale, rt):
    pos = jtp.to_fixed(w.pos, w.box_size)
    vel = (rng.standard_normal((w.n, 3)) * vscale).astype(rt)
    return pos, vel


@pytest.mark.parametrize("precision", [1, 0])
def test_dkd_matches_host_composition(jtp, precision):
    """DKD = host drift(dt/2) -> library force -> host kick(dt, friction) -> host drift(dt/2)."""
    print(f"case: DKD step vs host composition precision={precision}", flush=True)
    w = synth.clustered(n_total=20_000, box_size=100.0, n_cells=32, n_halos=20, seed=31)
    rt = np.float64 if precision == 1 else np.float32
    rng = np.random.default_rng(1)
    pos, vel = _state(jtp, w, rng, 0.5, rt)
    dt, H, dtau = 0.05, 0.07, 1.3
    ctx = _ctx(jtp, w, precision)
    n = w.n
    dp, dv = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(vel)
    ds, da = jtp.DeviceArray.from_host(w.species), jtp.DeviceArray((n, 3), rt)
    ctx.step_device(n, dp, dv, ds, da, dt, H, H, dtau, scheme=jtp.STEP_DKD)
    p_dev, v_dev, a_dev = dp.download(), dv.download(), da.download()
    # host composition
    p1 = _host_drift(pos, vel, 0.5 * dt, w.box_size)
    a_host = ctx.compute_forces(p1, w.species)
    assert np.array_equal(a_host, a_dev)  # the force at the drifted positions is bit-identical
    v64 = vel.astype(np.float64)
    v1 = (v64 + (a_host.astype(np.float64) - H * dtau * v64) * dt).astype(rt)
    ulp = np.finfo(rt).eps
    assert np.allclose(v_dev, v1, rtol=4 * ulp, atol=4 * ulp * np.abs(v1).max())
    p2 = _host_drift(p1, v_dev, 0.5 * dt, w.box_size)
    diff = (p_dev.astype(np.int64) - p2.astype(np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)
    assert np.abs(diff).max() == 0


@pytest.mark.parametrize("scheme", ["dkd", "kdk"])
def test_graph_replay_matches_eager(jtp, scheme):
    """A captured step replayed several times (with eager work and synchronisations in between, and dt / H
    changing between steps) gives bit-identical states to eager launches."""
    print(f"case: graph replay vs eager scheme={scheme}", flush=True)
    w = synth.clustered(n_total=30_000, box_size=100.0, n_cells=32, n_halos=20, seed=32)
    rng = np.random.default_rng(2)
    pos, vel = _state(jtp, w, rng, 0.5, np.float32)
    sch = jtp.STEP_DKD if scheme == "dkd" else jtp.STEP_KDK
    n = w.n
    ce, cg = _ctx(jtp, w, 0), _ctx(jtp, w, 0)
    de = [jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(vel), jtp.DeviceArray.from_host(w.species),
          jtp.DeviceArray((n, 3), np.float32)]
    dg = [jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(vel), jtp.DeviceArray.from_host(w.species),
          jtp.DeviceArray((n, 3), np.float32)]
    if sch == jtp.STEP_KDK:  # a(t) for the first half kick
        ce.compute_forces_device(n, de[0], de[2], de[3])
        cg.compute_forces_device(n, dg[0], dg[2], dg[3])
    for k, (dt, H) in enumerate([(0.02, 0.0), (0.03, 0.1), (0.01, 0.05), (0.02, 0.0)]):
        ce.step_device(n, de[0], de[1], de[2], de[3], dt, H, 0.5 * H, 1.0, scheme=sch, graph=False)
        cg.step_device(n, dg[0], dg[1], dg[2], dg[3], dt, H, 0.5 * H, 1.0, scheme=sch, graph=True)
        for a, b in zip(de, dg):
            assert np.array_equal(a.download(), b.download()), f"step {k}"
        st = cg.stats()
        assert st["ms_total"] > 0 and st["ms_walk"] == -1  # phases are not timed inside a graph
        # eager work on the same context between replays
        cg.compute_forces(pos, w.species)


def test_friction_kick_closed_form(jtp):
    """Zero masses: no force; DKD gives v_k = v0 (1 - H_s dtau dt)^k per species and positions advance by
    the trapezoid of the drifts."""
    print("case: friction closed form", flush=True)
    L = 100.0
    p = jtp.default_params(L, 32)
    p.precision = 1
    ctx = jtp.ForceContext(p, 0.0, 0.0)
    x = np.array([[1.0, 2.0, 3.0], [-10.0, 5.0, 20.0]])
    sp = np.array([1, -1], np.int8)
    v0 = np.array([[0.3, -0.2, 0.1], [-0.4, 0.25, 0.05]])
    pos = jtp.to_fixed(x, L)
    dp, dv = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(v0.copy())
    ds, da = jtp.DeviceArray.from_host(sp), jtp.DeviceArray((2, 3), np.float64)
    dt, Hp, Hm, dtau, K = 0.1, 0.2, 0.05, 1.5, 6
    for _ in range(K):
        ctx.step_device(2, dp, dv, ds, da, dt, Hp, Hm, dtau)
    v = dv.download()
    f = np.array([1 - Hp * dtau * dt, 1 - Hm * dtau * dt])[:, None]
    assert np.allclose(v, v0 * f ** K, rtol=1e-14)
    assert np.all(da.download() == 0)
    # displacement: sum over steps of dt/2 (v_{k} + v_{k+1})
    disp = sum(0.5 * dt * (v0 * f ** k + v0 * f ** (k + 1)) for k in range(K))
    got = jtp.from_fixed(dp.download(), L) - jtp.from_fixed(pos, L)
    got -= L * np.rint(got / L)
    assert np.allclose(got, disp, atol=4 * L / TWO32 * K)


def test_step_rejects_bad_arguments(jtp):
    print("case: step argument checks", flush=True)
    p = jtp.default_params(100.0, 32)
    ctx = jtp.ForceContext(p)
    lib = jtp.load_library()
    d = jtp.DeviceArray((4, 3), np.float32)
    assert lib.jtp_step(ctx.h, 4, d.ptr, d.ptr, d.ptr, d.ptr, 7, 0.1, 0, 0, 1, None) == jtp.JTP_ERR_ARG
    assert lib.jtp_step(ctx.h, 4, None, d.ptr, d.ptr, d.ptr, 0, 0.1, 0, 0, 1, None) == jtp.JTP_ERR_ARG
    assert lib.jtp_step(ctx.h, 4, d.ptr, d.ptr, d.ptr, d.ptr, 0, float("nan"), 0, 0, 1, None) == jtp.JTP_ERR_ARG
    assert lib.jtp_step(ctx.h, -1, d.ptr, d.ptr, d.ptr, d.ptr, 0, 0.1, 0, 0, 1, None) == jtp.JTP_ERR_ARG


@pytest.mark.parametrize("precision", [1, 0])
def test_two_body_energy(jtp, precision):
    """treepm_physics_8p.rs:267-374: two m+ particles at x = +-10 with v_y = +-0.1 in L = 100, G = 64,
    dt = 0.01, 100 leapfrog steps; E = KE - m^2/r must change by < 1 %.  Here the force is the full TreePM
    force and the integrator the library's KDK step."""
    print(f"case: two-body energy precision={precision}", flush=True)
    L, G = 100.0, 64
    p = jtp.default_params(L, G)
    p.precision = precision
    ctx = jtp.ForceContext(p, 1.0, 1.0)
    rt = np.float64 if precision == 1 else np.float32
    x = np.array([[10.0, 0, 0], [-10.0, 0, 0]])
    v = np.array([[0, 0.1, 0], [0, -0.1, 0]], rt)
    sp = np.array([1, 1], np.int8)
    dp, dv = jtp.DeviceArray.from_host(jtp.to_fixed(x, L)), jtp.DeviceArray.from_host(v)
    ds, da = jtp.DeviceArray.from_host(sp), jtp.DeviceArray((2, 3), rt)
    ctx.compute_forces_device(2, dp, ds, da)

    def energy():
        xx = jtp.from_fixed(dp.download(), L)
        vv = dv.download().astype(np.float64)
        d = xx[1] - xx[0]
        d -= L * np.rint(d / L)
        return 0.5 * np.sum(vv ** 2) - 1.0 / np.linalg.norm(d), vv

    e0, _ = energy()
    for _ in range(100):
        ctx.step_device(2, dp, dv, ds, da, 0.01, scheme=jtp.STEP_KDK)
    e1, vv = energy()
    de = abs(e1 - e0) / abs(e0)
    print(f"two-body dE/E = {de:.2e}")
    assert de < 1e-2
    assert np.abs(vv.sum(axis=0)).max() < 1e-5 * np.abs(vv).max()  # momentum


def test_kdk_reversibility_and_momentum(jtp):
    """FP64 KDK without friction is time-reversible: n steps, velocity flip, n steps, flip returns the start
    (to rounding).  With theta = 0 the short-range sum is pairwise antisymmetric and the PM force of a
    symmetric Green function, antisymmetric gradient and equal deposit / interpolation kernels has no
    self-force, so the total momentum of a one-species system is conserved to rounding."""
    print("case: KDK reversibility", flush=True)
    w = synth.clustered(n_total=8_000, box_size=100.0, n_cells=32, n_halos=10, seed=33)
    sp = np.ones(w.n, np.int8)
    L = w.box_size
    p = jtp.default_params(L, w.n_cells)
    p.r_s, p.r_cut, p.softening, p.precision, p.theta = w.r_s, w.r_cut, w.softening, 1, 0.0
    ctx = jtp.ForceContext(p, 1.0, 1.0)
    rng = np.random.default_rng(3)
    pos = jtp.to_fixed(w.pos, L)
    vel = rng.standard_normal((w.n, 3)) * 0.2
    vel -= vel.mean(axis=0)
    dp, dv = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(vel)
    ds, da = jtp.DeviceArray.from_host(sp), jtp.DeviceArray((w.n, 3), np.float64)
    ctx.compute_forces_device(w.n, dp, ds, da)
    p0 = vel.sum(axis=0)
    for _ in range(8):
        ctx.step_device(w.n, dp, dv, ds, da, 0.02, scheme=jtp.STEP_KDK, graph=True)
    mom = dv.download().sum(axis=0)
    print("momentum drift", np.abs(mom - p0).max(), "of", np.abs(vel).sum())
    assert np.abs(mom - p0).max() < 1e-9 * np.abs(vel).sum()
    dv.upload(-dv.download())
    ctx.compute_forces_device(w.n, dp, ds, da)
    for _ in range(8):
        ctx.step_device(w.n, dp, dv, ds, da, 0.02, scheme=jtp.STEP_KDK, graph=True)
    back = dp.download()
    diff = (back.astype(np.int64) - pos.astype(np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)
    print("reversibility: max |dx| =", np.abs(diff).max() * L / TWO32)
    assert np.abs(diff).max() * L / TWO32 < 1e-6 * L
    assert np.allclose(-dv.download(), vel, atol=1e-9)


@pytest.mark.parametrize("precision", [1, 0])
def test_diagnostics_match_host_sums(jtp, precision):
    """jtp_diagnostics: FP64 device reductions equal host FP64 sums of the same arrays (kinetic energy and
    momentum per species; 1/2 sum rho phi over both interlaced grids of the last force call)."""
    print(f"case: diagnostics precision={precision}", flush=True)
    w = synth.clustered(n_total=50_000, box_size=100.0, n_cells=32, n_halos=20, seed=34)
    rt = np.float64 if precision == 1 else np.float32
    rng = np.random.default_rng(4)
    vel = rng.standard_normal((w.n, 3)).astype(rt)
    ctx = _ctx(jtp, w, precision)
    pos = jtp.to_fixed(w.pos, w.box_size)
    ctx.compute_forces(pos, w.species)
    dv, ds = jtp.DeviceArray.from_host(vel), jtp.DeviceArray.from_host(w.species)
    d = ctx.diagnostics(w.n, dv, ds)
    v64 = vel.astype(np.float64)
    for k, (sg, m) in enumerate(((1, w.m_plus), (-1, w.m_minus_abs))):
        sel = w.species == sg
        ke = 0.5 * m * np.sum(v64[sel] ** 2)
        mom = m * v64[sel].sum(axis=0)
        assert abs(d["kinetic"][k] - ke) <= 1e-12 * ke
        assert np.allclose(d["momentum"][k], mom, rtol=1e-10, atol=1e-10 * m * np.abs(v64[sel]).sum())
    rho, phi = ctx.stage(jtp.STAGE_RHO).astype(np.float64), ctx.stage(jtp.STAGE_PHI).astype(np.float64)
    rho2, phi2 = ctx.stage(jtp.STAGE_RHO2).astype(np.float64), ctx.stage(jtp.STAGE_PHI2).astype(np.float64)
    e = 0.25 * (np.dot(rho, phi) + np.dot(rho2, phi2))
    assert abs(d["pm_energy"] - e) <= 1e-10 * abs(e)
    assert d["pm_energy"] != 0
