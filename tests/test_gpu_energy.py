"""FP64 potential energy (jtp_potential_energy) against lattice sums with closed forms:

* one mass alone in the periodic box (with the uniform background the mean-density removal
  implies): its energy with its own images is W = +1/2 G m^2 (2.837297479 / L), the simple cubic
  Madelung constant with the gravitational sign;
* one species on a simple cubic lattice of spacing a: W / N = +1/2 G m^2 (2.837297479 / a);
* two species on a simple cubic lattice, alternating like NaCl (every nearest neighbour of the
  other species): under the Janus sign rule the pair energy is -sigma_i sigma_j G m^2 / r, i.e. minus
  the electrostatic energy of charges sigma, so W / N = +1.747564595 G m^2 / (2 r0) with the NaCl
  Madelung constant and nearest-neighbour distance r0.

For such lattices the Ewald reciprocal sum vanishes (its terms carry exp(-(2 pi r_s / a)^2)), so the
Ewald energy of the TreePM force law (short-range pairs cut at r_cut) is, term by term, the
short-range lattice sum of -erfc(rho / 2 r_s) / rho inside r_cut, plus 1/2 N / (sqrt(pi) r_s) (the
long-range field of all the others, which on a lattice equals minus the smooth kernel's value at
r = 0), plus the background term 2 pi G r_s^2 Q^2 / V.  The test compares the library with that
decomposition (evaluated here in numpy) to 0.3 %, and with the Madelung closed form to within the
r_cut truncation of the short-range potential.  The residual is the mesh's error on the pairs at a
few cells (a CPU emulation of the mesh terms gives +0.05 % and +0.09 % at r_s = 2.5 cells, +0.5 % and
+0.8 % at 1.25 cells), so r_s = 2.5 cells here."""
import math

import numpy as np
import pytest
from scipy.special import erfc

pytestmark = pytest.mark.gpu


def _lattice(n_side, L, species_fn):
    g = (np.arange(n_side) + 0.25) * (L / n_side) - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    I, J, K = np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij")
    sp = species_fn(I, J, K).reshape(-1).astype(np.int8)
    return pos, sp


def _params(jtp, L, G):
    p = jtp.default_params(L, G)
    p.precision = 1
    p.r_s = 2.5 * L / G
    p.r_cut = 4.5 * p.r_s
    return p


def _energy(jtp, pos, sp, p):
    ctx = jtp.ForceContext(p, 1.0, 1.0)
    dp = jtp.DeviceArray.from_host(jtp.to_fixed(pos, p.box_size))
    ds = jtp.DeviceArray.from_host(sp)
    return ctx.potential_energy(len(sp), dp, ds)


def _decomposition(p, a, n_side, alternating):
    """TreePM energy of the lattice from its terms (see the module docstring)."""
    R = int(math.ceil(p.r_cut / a)) + 1
    ii = np.arange(-R, R + 1)
    X, Y, Z = np.meshgrid(ii, ii, ii, indexing="ij")
    r = a * np.sqrt(X ** 2 + Y ** 2 + Z ** 2).ravel()
    sg = np.where((X + Y + Z) % 2 == 0, 1.0, -1.0).ravel() if alternating else np.ones(r.size)
    m = (r > 0) & (r < p.r_cut)
    rho = np.sqrt(r[m] ** 2 + p.softening ** 2)
    s_short = np.sum(sg[m] * erfc(rho / (2 * p.r_s)) / rho)
    N = n_side ** 3
    Q = 0.0 if alternating else float(N)
    return -0.5 * N * s_short + 0.5 * N / (math.sqrt(math.pi) * p.r_s) + \
        2 * math.pi * p.r_s ** 2 * Q * Q / p.box_size ** 3


def test_single_mass_images(jtp):
    print("case: energy single mass", flush=True)
    L, G = 100.0, 64
    p = _params(jtp, L, G)
    for sp, m in ((1, 1.0), (-1, 2.5)):
        ctx = jtp.ForceContext(p, m if sp > 0 else 1.0, m if sp < 0 else 1.0)
        dp = jtp.DeviceArray.from_host(jtp.to_fixed(np.array([[3.3, -7.1, 12.9]]), L))
        ds = jtp.DeviceArray.from_host(np.array([sp], np.int8))
        w = ctx.potential_energy(1, dp, ds)
        expect = 0.5 * m * m * 2.837297479 / L
        print("species %d, m = %g: W = %.10g, expected %.10g" % (sp, m, w, expect))
        assert abs(w / expect - 1) < 1e-9


def test_simple_cubic_madelung(jtp):
    print("case: energy simple cubic", flush=True)
    L, n_side, G = 100.0, 32, 64
    pos, sp = _lattice(n_side, L, lambda i, j, k: np.ones_like(i))
    p = _params(jtp, L, G)
    w = _energy(jtp, pos, sp, p)
    a = L / n_side
    expect = 0.5 * len(sp) * (2.837297479 / a)
    dec = _decomposition(p, a, n_side, False)
    print("W = %.8g, decomposition %.8g (rel %.2e), Madelung %.8g (rel %.2e)"
          % (w, dec, w / dec - 1, expect, w / expect - 1))
    assert abs(w / dec - 1) < 3e-3
    assert abs(w / expect - 1) < abs(dec / expect - 1) + 3e-3


def test_nacl_janus_madelung(jtp):
    print("case: energy NaCl", flush=True)
    L, n_side, G = 100.0, 32, 64
    pos, sp = _lattice(n_side, L, lambda i, j, k: np.where((i + j + k) % 2 == 0, 1, -1))
    p = _params(jtp, L, G)
    w = _energy(jtp, pos, sp, p)
    r0 = L / n_side
    expect = len(sp) * 1.747564595 / (2 * r0)
    dec = _decomposition(p, r0, n_side, True)
    print("W = %.8g, decomposition %.8g (rel %.2e), Madelung %.8g (rel %.2e)"
          % (w, dec, w / dec - 1, expect, w / expect - 1))
    assert abs(w / dec - 1) < 3e-3
    assert abs(w / expect - 1) < abs(dec / expect - 1) + 3e-3


def test_energy_rejects_fp32(jtp):
    print("case: energy argument checks", flush=True)
    p = jtp.default_params(100.0, 32)
    ctx = jtp.ForceContext(p)
    d = jtp.DeviceArray((8, 3), np.uint32)
    s = jtp.DeviceArray((8,), np.int8)
    with pytest.raises(jtp.JtpError):
        ctx.potential_energy(8, d, s)
