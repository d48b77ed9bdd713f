"""The reference's qualitative TreePM tests, run on the GPU library:

* force continuity across r_cut (tests/treepm_continuity.rs:34-78): adjacent-radius jumps < 10 %,
  and here also the force itself against the FP64 Ewald sum;
* PM isotropy (tests/treepm_isotropy.rs:52-224): the force of a central mass on probes on a
  Fibonacci sphere points radially (sigma and mean of the angle < 5 deg at 64^3, sigma < 3 deg at 128^3),
  for attraction (+ on +, - on -) and repulsion (+ on -);
* Gaussian splitting (src/treepm/pm_grid.rs:373-416): the split PM force is < 0.5 of the unsplit one
  at r = 5 and > 0.8 of it at r = 40 (r_s = 10, L = 100, G = 64).

The reference measures the field of a mass at a probe that carries no mass.  The library takes one
mass per species, so the probes here are particles and the central mass's field is isolated by
linearity: a(centre + probes) - a(probes), with the PM-only FP64 form (exactly linear up to rounding).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fibonacci_sphere(n, radius):
    """treepm_isotropy.rs:10-24."""
    phi = np.pi * (3.0 - np.sqrt(5.0))
    i = np.arange(n)
    y = 1.0 - i / (n - 1) * 2.0
    r = np.sqrt(1.0 - y * y)
    th = phi * i
    return np.stack([radius * r * np.cos(th), radius * y, radius * r * np.sin(th)], axis=1)


def _pm_field(jtp, L, G, centre_sp, centre_mass, probes, probe_sp, **kw):
    """Acceleration of the probes due to one mass at the origin (PM only, FP64), by linearity."""
    p = jtp.default_params(L, G)
    p.tree_enabled, p.precision = 0, 1
    for k, v in kw.items():
        setattr(p, k, v)
    m_plus = centre_mass if centre_sp > 0 else 1.0
    m_minus = centre_mass if centre_sp < 0 else 1.0  # probes of the centre's species share its mass
    ctx = jtp.ForceContext(p, m_plus, m_minus)
    xp = np.concatenate([np.zeros((1, 3)), probes])
    sp = np.concatenate([[centre_sp], np.full(len(probes), probe_sp)]).astype(np.int8)
    a_with = ctx.compute_forces(jtp.to_fixed(xp, L), sp)[1:]
    a_without = ctx.compute_forces(jtp.to_fixed(probes, L), sp[1:])
    return a_with - a_without


def _angles(pos, force, outward):
    r = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    if not outward:
        r = -r
    f = force / np.linalg.norm(force, axis=1, keepdims=True)
    return np.degrees(np.arccos(np.clip(np.sum(f * r, axis=1), -1, 1)))


@pytest.mark.parametrize("centre_sp,probe_sp,G,radius,n,bound", [
    (1, 1, 64, 30.0, 100, 5.0),     # test_pm_isotropy_positive
    (-1, -1, 64, 30.0, 100, 5.0),   # test_pm_isotropy_negative
    (1, -1, 64, 30.0, 100, 5.0),    # test_pm_repulsion_isotropy
    (1, 1, 128, 40.0, 200, 3.0),    # test_pm_isotropy_fine_grid
])
def test_pm_isotropy(jtp, centre_sp, probe_sp, G, radius, n, bound):
    print(f"case: PM isotropy centre={centre_sp} probe={probe_sp} G={G}", flush=True)
    probes = fibonacci_sphere(n, radius)
    a = _pm_field(jtp, 200.0, G, centre_sp, 100.0, probes, probe_sp)
    ang = _angles(probes, a, outward=(centre_sp != probe_sp))
    print("angle from radial: mean %.3f deg, sigma %.3f deg, max %.3f deg" % (ang.mean(), ang.std(), ang.max()))
    assert ang.std() < bound and ang.mean() < bound
    # magnitude: the field of the central mass at the shell, within the periodic correction
    mag = np.linalg.norm(a, axis=1)
    assert np.all(np.abs(mag / (100.0 / radius ** 2) - 1.0) < 0.2)


def test_gaussian_split_reduces_short_range(jtp):
    print("case: Gaussian split", flush=True)
    L, G = 100.0, 64
    probes = np.array([[5.0, 0, 0], [40.0, 0, 0]])
    full = _pm_field(jtp, L, G, 1, 1.0, probes, 1, r_s=0.0, split_mode=2, deconvolve=1)
    split = _pm_field(jtp, L, G, 1, 1.0, probes, 1, r_s=10.0)
    ratio = split[:, 0] / full[:, 0]
    print("split / full at r = 5, 40:", ratio)
    assert abs(split[0, 0]) < 0.5 * abs(full[0, 0])
    assert ratio[1] > 0.8


@pytest.mark.parametrize("precision", [1, 0])
def test_force_continuity_at_r_cut(jtp, orc, precision):
    """Probe at r = r_cut + i * 0.5 (i = -10..10) from a unit mass, r_cut = 20, L = 100, G = 64 (full
    TreePM: tree + PM).  Relative jumps between neighbours < 10 % as in the reference; the force itself
    within 0.3 % of the periodic Ewald force of the pair inside r_cut, and beyond r_cut within the
    short-range share the hard cut drops, S(r_cut / 2 r_s) = erfc(2.25) + 4.5/sqrt(pi) exp(-5.0625)
    = 1.76 % (plus 0.3 %)."""
    print(f"case: continuity precision={precision}", flush=True)
    L, G, r_cut = 100.0, 64, 20.0
    p = jtp.default_params(L, G)
    p.r_cut, p.r_s, p.precision, p.softening = r_cut, r_cut / 4.5, precision, 0.1
    ctx = jtp.ForceContext(p, 1.0, 1.0)
    radii = r_cut + 0.5 * np.arange(-10, 11)
    f, ref = [], []
    for r in radii:
        x = np.array([[0.0, 0, 0], [r, 0, 0]])
        pos = jtp.to_fixed(x, L)
        sp = np.array([1, 1], np.int8)
        a = ctx.compute_forces(pos, sp).astype(np.float64)
        f.append(a[1, 0])
        ref.append(orc.ewald(pos, sp, 1.0, 1.0, L, 0.1, targets=np.array([1]), rc=L / 3)[0, 0])
    f, ref = np.array(f), np.array(ref)
    jump = np.abs(np.diff(f)) / np.abs(f[:-1])
    err = np.abs(f / ref - 1)
    print("max jump %.3f" % jump.max())
    for r, e in zip(radii, err):
        print("  r = %5.1f  |F/F_ewald - 1| = %.2e" % (r, e))
    assert jump.max() < 0.10
    x = r_cut / (2 * p.r_s)
    s_cut = math.erfc(x) + 2 * x / math.sqrt(math.pi) * math.exp(-x * x)
    assert err[radii < r_cut].max() < 3e-3
    assert err[radii >= r_cut].max() < s_cut + 3e-3
