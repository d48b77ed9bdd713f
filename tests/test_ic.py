"""Zel'dovich initial conditions (jtp_generate_ics; reference generate_zeldovich_ics, src/ic_gen.rs).

The reference draws its field from StdRng (ChaCha12) and rand_distr's Normal, neither of which is
available here, so the library and the oracle share SplitMix64 streams instead and parity is: the
GPU path against the oracle's step-by-step restatement (full complex spectrum, complex inverse FFT,
real part / n^3) on the same streams.  The oracle itself is pinned by SplitMix64's published output
for seed 1234567, by the reference's own tests (src/ic_gen.rs:320-367: the 1:19 species ratio and
positions inside the box) and by what the construction fixes: max displacement = 0.3 spacing,
velocity = displacement x H0 sqrt(1 + z), a curl-free displacement field, and its power spectrum
slope n_s - 6."""
import math

import numpy as np
import pytest


def test_splitmix64_published_vector(orc):
    # Vigna's splitmix64.c, seed 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert orc.splitmix64(1234567, 5).tolist() == want


def test_reference_ic_tests(orc):
    """test_ic_generation_small / test_positions_in_box (ic_gen.rs:324-366), on the single-grid code
    path those tests exercise (one particle per cell: n^3 particles)."""
    L, n = 50.0, 4
    pos, vel, sign, n_plus = orc.ic(n, L, z_init=10.0, seed=42, mu=19.0, n_s=0.965, delta_rms=0.10)
    assert len(sign) == n ** 3 and pos.shape == (n ** 3, 3) and vel.shape == (n ** 3, 3)
    n_minus = int(np.sum(sign < 0))
    assert n_plus == int(np.sum(sign > 0)) == round(n ** 3 / 20)
    assert abs(n_minus / n_plus - 19.0) < 2.0
    assert np.all(np.abs(pos) <= L / 2)


def _displacement(pos, n, L):
    h = L / n
    g = (np.arange(n) + 0.5) * h - L / 2
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([X, Y, Z], -1).reshape(-1, 3)
    d = pos - lat
    return d - L * np.round(d / L)


def test_oracle_construction_invariants(orc):
    L, n, z = 100.0, 32, 10.0
    pos, vel, sign, n_plus = orc.ic(n, L, z_init=z, seed=7)
    d = _displacement(pos, n, L)
    assert np.max(np.linalg.norm(d, axis=1)) == pytest.approx(0.3 * L / n, rel=1e-12)
    vs = 69.9 / 977.8 * math.sqrt(1 + z)
    assert np.allclose(vel, d * vs, rtol=1e-9, atol=1e-12 * vs)
    # curl-free: k x psi(k) = 0 away from the Nyquist planes
    D = [np.fft.fftn(d[:, a].reshape(n, n, n)) for a in range(3)]  # axes (z, y, x)
    k = np.fft.fftfreq(n) * n
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    ok = (np.abs(KX) < n / 2) & (np.abs(KY) < n / 2) & (np.abs(KZ) < n / 2) & ((KX != 0) | (KY != 0) | (KZ != 0))
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    mag = np.sqrt(sum(np.abs(x) ** 2 for x in D))
    curl = np.abs(KX * D[1] - KY * D[0]) + np.abs(KY * D[2] - KZ * D[1]) + np.abs(KZ * D[0] - KX * D[2])
    assert np.max(curl[ok] / (kk[ok] * mag[ok] + 1e-300)) < 1e-9
    # displacement power ~ k^(n_s - 6): slope of shell means between 2 and 8 fundamental modes
    p = (mag ** 2)[ok]
    kb = kk[ok]
    shells = np.arange(2, 9)
    means = [p[(kb >= s - 0.5) & (kb < s + 0.5)].mean() for s in shells]
    slope = np.polyfit(np.log(shells), np.log(means), 1)[0]
    assert slope == pytest.approx(0.965 - 6, abs=0.3)
    # the sign set: n+ cells with the smallest SplitMix64(seed + 12345) keys
    keys = orc.splitmix64(7 + 12345, n ** 3)
    plus = np.zeros(n ** 3, bool)
    plus[np.lexsort((np.arange(n ** 3), keys))[:n_plus]] = True
    assert np.array_equal(sign > 0, plus)


def _fixed_diff(a, b):
    return np.abs(((a.astype(np.int64) - b.astype(np.int64) + 2 ** 31) % 2 ** 32) - 2 ** 31)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_gpu_ics_match_oracle(jtp, orc, precision):
    print(f"case: ICs vs oracle precision={precision}", flush=True)
    L, n = 100.0, 32
    p = jtp.default_params(L, n)
    p.precision = precision
    ctx = jtp.ForceContext(p, 1.0, 19.0)
    ip = jtp.ic_params(box_size=L, n_grid=n, seed=7)
    dp = jtp.DeviceArray((n ** 3, 3), np.uint32)
    dv = jtp.DeviceArray((n ** 3, 3), np.float64)
    ds = jtp.DeviceArray((n ** 3,), np.int8)
    n_plus = ctx.generate_ics(ip, dp, dv, ds, vel_f64=True)
    pos, vel, sign, n_plus_ref = orc.ic(n, L, z_init=ip.z_init, seed=7, mu=ip.mu, n_s=ip.n_s, delta_rms=ip.delta_rms)
    assert n_plus == n_plus_ref
    assert np.array_equal(ds.download(), sign)
    h_units = 2 ** 32 / n  # one lattice spacing in fixed-point units
    tol = 4 if precision == 1 else 1e-5 * h_units
    assert _fixed_diff(dp.download(), orc.fixed(pos, L)).max() <= tol
    rtol = 1e-10 if precision == 1 else 1e-5
    v = dv.download()
    assert np.max(np.abs(v - vel)) <= rtol * np.max(np.abs(vel))
    # a second call gives the same bits
    dp2 = jtp.DeviceArray((n ** 3, 3), np.uint32)
    ctx.generate_ics(ip, dp2, dv, ds, vel_f64=True)
    assert np.array_equal(dp2.download(), dp.download())
    # the ICs drive a force evaluation
    acc = ctx.compute_forces(dp.download(), ds.download())
    assert np.all(np.isfinite(acc))


@pytest.mark.gpu
def test_gpu_ics_rejects_mismatch(jtp):
    print("case: IC argument checks", flush=True)
    ctx = jtp.ForceContext(jtp.default_params(100.0, 32))
    d = jtp.DeviceArray((64 ** 3, 3), np.uint32)
    s = jtp.DeviceArray((64 ** 3,), np.int8)
    with pytest.raises(jtp.JtpError):
        ctx.generate_ics(jtp.ic_params(box_size=100.0, n_grid=64), d, d, s)
    with pytest.raises(jtp.JtpError):
        ctx.generate_ics(jtp.ic_params(box_size=50.0, n_grid=32), d, d, s)
