"""CPU checks that pin the oracle before it is trusted: numpy golden vectors,
closed forms, brute force on tiny inputs, and the reference's own qualitative
tests (treepm_force.rs:152-266, tree_short.rs:254-355, pm_grid.rs:373-460,
treepm_isotropy.rs:52-133) re-run on the restated reference CPU path."""
import os

import numpy as np
import pytest
from scipy.special import erfc

from oracle.oracle_lib import RefCpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_fft_poisson.npz")


def interleave_np(pos):
    # independent Morton interleave: bit b of axis a -> bit 3b + a
    key = np.zeros(len(pos), np.uint64)
    for a in range(3):
        v = (pos[:, a] >> np.uint32(11)).astype(np.uint64)
        for b in range(21):
            key |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return key


def test_morton_fixed(orc):
    rng = np.random.default_rng(0)
    pos = rng.integers(0, 2 ** 32, (5000, 3), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(orc.morton(pos), interleave_np(pos))
    k = orc.morton(np.array([[2 ** 31, 0, 0], [0, 2 ** 31, 0], [0, 0, 2 ** 31]], np.uint32))
    assert list(k) == [1 << 60, 1 << 61, 1 << 62]


def test_morton_f32_reference_formula(orc):
    L = 250.0
    bh, inv = np.float32(L / 2), np.float32(2097152.0 / L)
    rng = np.random.default_rng(1)
    x = rng.uniform(-L / 2, L / 2, (4000, 3)).astype(np.float32)
    x[0] = -L / 2
    x[1] = L / 2
    x[2] = -200.0  # clamps to 0
    g = ((x + bh) * inv).astype(np.float32)  # IEEE single add then multiply
    c = np.clip(g, np.float32(0), np.float32(2097151)).astype(np.uint32)
    want = np.zeros(len(x), np.uint64)
    for a in range(3):
        v = c[:, a].astype(np.uint64)
        for b in range(21):
            want |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    got = orc.morton_f32_ref(x, bh, inv)
    assert np.array_equal(got, want)
    assert got[1] == (1 << 63) - 1 and got[2] == 0


def test_morton30_reference_formula(orc):
    """compute_morton_f32 (nbody_gpu_twopass.rs:302-326) evaluated independently in numpy IEEE single:
    scale = 1024 / (2 box_half), saturating conversion, min 1023, 10-bit interleave, key << 32 | index."""
    L = 250.0
    bh = np.float32(L / 2)
    rng = np.random.default_rng(11)
    x = rng.uniform(-L / 2, L / 2, (5000, 3)).astype(np.float32)
    x[0] = -L / 2
    x[1] = L / 2            # 1024 * scale rounds to the top cell, clamps to 1023
    x[2] = -200.0           # negative before the conversion: 0
    x[3] = 1e30             # beyond 2^32: saturates, then 1023
    x[4] = (0.5, 0.25, -0.125)
    scale = np.float32(1024.0) / (np.float32(2.0) * bh)
    g = ((x + bh) * scale).astype(np.float32)
    c = np.where(g > 0, np.minimum(np.where(g >= np.float32(2 ** 32), np.float32(2 ** 32 - 1), g), 1023.0), 0)
    c = c.astype(np.uint64)
    mc = np.zeros(len(x), np.uint64)
    for a in range(3):
        for b in range(10):
            mc |= ((c[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    want = (mc << np.uint64(32)) | np.arange(len(x), dtype=np.uint64)
    keys, idx = orc.morton30_f32_ref(x, bh)
    assert np.array_equal(keys, want)
    assert np.array_equal(idx, np.arange(len(x)))
    assert keys[1] >> np.uint64(32) == (1 << 30) - 1 and keys[2] >> np.uint64(32) == 0
    assert keys[3] >> np.uint64(32) == (1 << 30) - 1
    # a known answer: cells (514, 513, 511) of x = (0.5, 0.25, -0.125) at L = 250 (cell 0.244 Mpc)
    cx, cy, cz = 514, 513, 511
    sp = lambda v: sum(((v >> b) & 1) << (3 * b) for b in range(10))
    assert int(keys[4]) == ((sp(cx) | (sp(cy) << 1) | (sp(cz) << 2)) << 32) | 4


def test_radix_sort_is_stable_argsort(orc):
    rng = np.random.default_rng(2)
    k = rng.integers(0, 2 ** 63, 30000, dtype=np.uint64)
    k[rng.integers(0, 30000, 8000)] = k[rng.integers(0, 30000, 8000)]
    k[:100] = 7
    assert np.array_equal(orc.radix_sort(k), np.argsort(k, kind="stable"))


def test_karras_bruteforce(orc):
    rng = np.random.default_rng(3)
    for n in (2, 3, 17, 64, 200):
        k = np.sort(rng.integers(0, 2 ** 20, n, dtype=np.uint64))  # many duplicates for n = 200
        nodes = orc.karras(k)
        ext = lambda i: (int(k[i]) << 64) | i
        seen = []

        def check(node, f, l):
            seen.append(node)
            assert (nodes[node, 0], nodes[node, 1]) == (f, l)
            hb = (ext(f) ^ ext(l)).bit_length() - 1
            s = max(i for i in range(f, l + 1) if ((ext(i) >> hb) & 1) == 0)
            for child, cf, cl in ((nodes[node, 2], f, s), (nodes[node, 3], s + 1, l)):
                if cf == cl:
                    assert child == ~cf
                else:
                    check(child, cf, cl)

        check(0, 0, n - 1)
        assert sorted(seen) == list(range(n - 1))


def test_node_sums_prefix(orc):
    rng = np.random.default_rng(4)
    n = 3000
    pos = rng.integers(0, 2 ** 32, (n, 3), dtype=np.uint64).astype(np.uint32)
    sp = np.where(rng.random(n) < 0.3, 1, -1).astype(np.int8)
    keys = orc.morton(pos)
    perm = np.argsort(keys, kind="stable")
    nodes = orc.karras(keys[perm])
    sums, boxes = orc.node_sums(pos[perm], sp[perm], nodes)
    P = pos[perm].astype(np.uint64)
    plus = (sp[perm] > 0).astype(np.uint64)
    cp = np.concatenate([[0], np.cumsum(plus)])
    for a in range(3):
        c = np.concatenate([[0], np.cumsum(P[:, a] * plus)])
        assert np.array_equal(sums[:, 1 + a], c[nodes[:, 1] + 1] - c[nodes[:, 0]])
    assert np.array_equal(sums[:, 0], cp[nodes[:, 1] + 1] - cp[nodes[:, 0]])
    assert np.array_equal(sums[:, 0] + sums[:, 4], (nodes[:, 1] - nodes[:, 0] + 1).astype(np.uint64))
    i = 5
    seg = pos[perm][nodes[i, 0]:nodes[i, 1] + 1]
    assert np.array_equal(boxes[i, :3], seg.min(0)) and np.array_equal(boxes[i, 3:], seg.max(0))


def test_fft_golden(orc):
    g = np.load(GOLD)
    for G in (8, 16):
        out = orc.fft3d(g[f"fft_in_{G}"])
        assert np.allclose(out, g[f"fft_out_{G}"], rtol=0, atol=1e-11 * np.abs(g[f"fft_out_{G}"]).max())
        back = orc.fft3d(out, inverse=True) / G ** 3
        assert np.allclose(back, g[f"fft_in_{G}"], atol=1e-12)


def test_poisson_golden(orc):
    g = np.load(GOLD)
    phi = orc.poisson(g["poisson_rho"], 16, 40.0, 1.0, 2.0, 0, 1).reshape(16, 16, 16)
    ref = g["poisson_phi"]
    assert np.max(np.abs(phi - ref)) < 1e-12 * np.max(np.abs(ref))


def test_poisson_matched_stencil(orc):
    # a single mode along x: the matched factor is k / D(k), so the stencil gradient of phi is exact
    G, L = 16, 10.0
    h = L / G
    x = np.arange(G) * h
    m = 3
    rho = np.broadcast_to(np.cos(2 * np.pi * m * x / L), (G, G, G)).copy()
    k = 2 * np.pi * m / L
    D4 = (8 * np.sin(k * h) - np.sin(2 * k * h)) / (6 * h)
    phi = orc.poisson(rho, G, L, 1.0, 0.0, 0, 0, match=0).reshape(G, G, G)
    assert np.allclose(phi, -4 * np.pi / k ** 2 * (k / D4) * rho / h ** 3, atol=1e-11)
    D2 = np.sin(k * h) / h
    phi = orc.poisson(rho, G, L, 1.0, 0.0, 0, 0, match=1).reshape(G, G, G)
    assert np.allclose(phi, -4 * np.pi / k ** 2 * (k / D2) * rho / h ** 3, atol=1e-11)


def test_poisson_single_mode_kmin(orc):
    # rho = cos(2 pi m x / L): phi = green(k) rho exactly; k_min above m removes it
    G, L = 16, 10.0
    x = np.arange(G) * L / G
    rho = np.broadcast_to(np.cos(2 * np.pi * 2 * x / L), (G, G, G)).copy()
    phi = orc.poisson(rho, G, L, 1.0, 0.0, 0, 0).reshape(G, G, G)
    k = 2 * np.pi * 2 / L
    assert np.allclose(phi, -4 * np.pi / k ** 2 * rho / (L / G) ** 3, atol=1e-12)
    phi = orc.poisson(rho, G, L, 1.0, 0.0, 3, 0)
    assert np.max(np.abs(phi)) < 1e-14


def test_deposit_closed_form(orc):
    G, L = 8, 8.0
    u = np.array([[int(2.25 / L * 2 ** 32), int(7.5 / L * 2 ** 32), int(0.125 / L * 2 ** 32)]], np.uint32)
    rho = orc.deposit(u, np.array([1], np.int8), 2.0, 0.0, G).reshape(G, G, G)
    fx, fy, fz = 0.25, 0.5, 0.125
    assert abs(rho.sum() - 2.0) < 1e-14
    assert abs(rho[0, 7, 2] - 2 * (1 - fx) * (1 - fy) * (1 - fz)) < 1e-14
    assert abs(rho[1, 0, 3] - 2 * fx * fy * fz) < 1e-14  # x+1, y wraps 7 -> 0, z+1
    rng = np.random.default_rng(5)
    pos = rng.integers(0, 2 ** 32, (10000, 3), dtype=np.uint64).astype(np.uint32)
    sp = np.where(rng.random(10000) < 0.5, 1, -1).astype(np.int8)
    rho = orc.deposit(pos, sp, 1.0, 19.0, 16)
    assert abs(rho.sum() - (np.sum(sp > 0) - 19.0 * np.sum(sp < 0))) < 1e-9


@pytest.mark.parametrize("four", [True, False])
def test_interp_sinusoid(orc, four):
    G, L, m = 16, 16.0, 3
    h = L / G
    x = np.arange(G) * h
    phi = np.broadcast_to(np.sin(2 * np.pi * m * x / L), (G, G, G)).copy()
    kh = 2 * np.pi * m / L * h
    d = (8 * np.sin(kh) - np.sin(2 * kh)) / (6 * h) if four else np.sin(kh) / h
    idx = np.arange(G)
    pos = np.stack([idx * (2 ** 32 // G), np.full(G, 5 * 2 ** 28), np.full(G, 2 ** 27)], 1).astype(np.uint32)
    acc = orc.interp(pos, np.ones(G, np.int8), G, L, phi, four)
    assert np.allclose(acc[:, 0], -d * np.cos(2 * np.pi * m * x / L), atol=1e-12)
    assert np.allclose(acc[:, 1:], 0, atol=1e-12)
    acc_m = orc.interp(pos, -np.ones(G, np.int8), G, L, phi, four)
    assert np.allclose(acc_m, -acc, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_short_direct_pair(orc, mode):
    L, eps, rcut, rs = 100.0, 0.3, 10.0, 2.0
    r = 3.0
    x = np.array([[0.0, 0, 0], [r, 0, 0]])
    u = orc.fixed(x, L)
    r = float(np.int64(u[1, 0]) - np.int64(u[0, 0])) * L / 2 ** 32  # quantised separation
    for si, sj in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
        sp = np.array([si, sj], np.int8)
        a = orc.short_direct(u, sp, 1.0, 19.0, L, eps, rcut, rs, mode)
        rp = np.sqrt(r * r + eps * eps)
        xx = rp / (2 * rs)
        S = {0: erfc(xx) + 2 * xx / np.sqrt(np.pi) * np.exp(-xx * xx), 1: erfc(xx), 2: 1 - (r / rcut) ** 4}[mode]
        m_j = 1.0 if sj > 0 else 19.0
        want = si * sj * m_j * S * r / rp ** 3
        assert abs(a[0, 0] - want) < 1e-9 * abs(want)
    a = orc.short_direct(orc.fixed(np.array([[0.0, 0, 0], [12.0, 0, 0]]), L), np.array([1, 1], np.int8), 1, 1, L,
                         eps, rcut, rs, mode)
    assert np.all(a == 0)


def test_short_direct_celllist_consistent(orc):
    rng = np.random.default_rng(6)
    n, L = 3000, 50.0
    u = rng.integers(0, 2 ** 32, (n, 3), dtype=np.uint64).astype(np.uint32)
    sp = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int8)
    a1 = orc.short_direct(u, sp, 1.0, 3.0, L, 0.1, 9.9, 2.0, 0)       # 5^3 cells
    a2 = orc.short_direct(u, sp, 1.0, 3.0, L, 0.1, 9.9, 2.0, 0, targets=np.arange(n))
    assert np.allclose(a1, a2, rtol=1e-14, atol=0)
    # brute force over all pairs in numpy
    X = (u.astype(np.float64) / 2 ** 32) * L
    i = 17
    d = X - X[i]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(1)
    ok = (r2 < 9.9 ** 2)
    r2s = r2 + 0.01
    xx = np.sqrt(r2s) / 4.0
    S = erfc(xx) + 2 * xx / np.sqrt(np.pi) * np.exp(-xx ** 2)
    q = np.where(sp > 0, 1.0, -3.0)
    f = np.where(ok, q * S / r2s ** 1.5, 0.0)
    want = sp[i] * (f[:, None] * d).sum(0)
    assert np.allclose(a1[i], want, rtol=1e-9)


def test_ewald_madelung(orc):
    L = 3.0
    for alpha in (1.0, 1.5):
        v = orc.lib.orc_ewald_sc_potential(L, alpha, 6, 12)
        assert abs(v * L - (-2.837297479)) < 1e-8


def test_ewald_alpha_independent_and_symmetric(orc):
    rng = np.random.default_rng(7)
    L, n = 10.0, 60
    u = rng.integers(0, 2 ** 32, (n, 3), dtype=np.uint64).astype(np.uint32)
    sp = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int8)
    # erfc(alpha rc) and exp(-k_max^2 / 4 alpha^2) both below 1e-12 for every choice; eps = 0 because the
    # Plummer correction of the nearest image is applied inside the real-space cutoff only
    a1 = orc.ewald_custom(u, sp, 1.0, 5.0, L, 0.0, alpha=1.0, rc=5.0, nk=18)
    a2 = orc.ewald_custom(u, sp, 1.0, 5.0, L, 0.0, alpha=1.2, rc=5.0, nk=21)
    a3 = orc.ewald_custom(u, sp, 1.0, 5.0, L, 0.0, alpha=0.8, rc=12.0, nk=15)  # includes real-space images
    scale = np.abs(a1).max()
    assert np.max(np.abs(a1 - a2)) < 1e-8 * scale
    assert np.max(np.abs(a1 - a3)) < 1e-8 * scale
    # two like particles half a box apart: zero force by symmetry
    u2 = orc.fixed(np.array([[0.0, 0, 0], [L / 2, 0, 0]]), L)
    a = orc.ewald_custom(u2, np.array([1, 1], np.int8), 1.0, 1.0, L, 0.0, alpha=1.0, rc=5.0, nk=18)
    assert np.max(np.abs(a)) < 1e-10
    # a close pair in a big box: Newton plus a small periodic correction
    u3 = orc.fixed(np.array([[0.0, 0, 0], [1.0, 0, 0]]), 1000.0)
    a = orc.ewald(u3, np.array([1, -1], np.int8), 1.0, 1.0, 1000.0, 0.0)
    assert abs(a[0, 0] - (-1.0)) < 1e-6 and abs(a[1, 0] - 1.0) < 1e-6  # the pair repels
    # with softening the close pair follows Plummer: r / (r^2 + eps^2)^{3/2}
    a = orc.ewald(u3, np.array([1, 1], np.int8), 1.0, 1.0, 1000.0, 0.5)
    assert abs(a[0, 0] - 1.0 / 1.25 ** 1.5) < 1e-6


# ---------------------------------------------------------------- reference tests on the restated CPU path
def _ref(orc, r_cut=20.0, G=32, L=100.0, theta=0.5, eps=0.5):
    return RefCpu(orc, r_cut, G, L, theta, eps, r_cut / 3.0, 1.0, 1.0)


def test_ref_treepm_basic(orc):
    r = _ref(orc)
    r.update(np.array([[-10.0, 0, 0], [10.0, 0, 0]]), np.ones(2), np.array([1, 1], np.int8))
    a = r.forces()
    assert a[0, 0] > 0 and a[1, 0] < 0
    assert abs(a[0, 0] + a[1, 0]) < 0.01 * abs(a[0, 0])


@pytest.mark.parametrize("si,sj,attract", [(1, 1, True), (-1, -1, True), (1, -1, False), (-1, 1, False)])
def test_ref_four_signs(orc, si, sj, attract):
    r = _ref(orc, r_cut=30.0)
    r.update(np.array([[-10.0, 0, 0], [10.0, 0, 0]]), np.ones(2), np.array([si, sj], np.int8))
    assert (r.forces()[0, 0] > 0) == attract


def test_ref_short_range_cutoff(orc):
    r = RefCpu(orc, 10.0, 32, 100.0, 0.5, 0.1, 10.0 / 3, 1.0, 1.0)
    src = np.array([[0.0, 0, 0]])
    for x, check in ((3.0, lambda v: v < -0.01), (8.0, lambda v: v < 0), (15.0, lambda v: abs(v) < 1e-10)):
        r.update(np.array([[0.0, 0, 0], [x, 0, 0]]), np.array([1.0, 0.0]), np.array([1, 1], np.int8), pm=False)
        a = r.forces(targets=np.array([1]))
        assert check(a[0, 0])


def test_ref_pm_gaussian_split_and_signs(orc):
    # pm_grid.rs:373-416: Gaussian damping reduces the short-range PM force by > 50 % at r=5, keeps > 0.8 at r=40
    full = RefCpu(orc, 30.0, 64, 100.0, 0.5, 0.5, 0.0, 1.0, 1.0)
    split = RefCpu(orc, 30.0, 64, 100.0, 0.5, 0.5, 10.0, 1.0, 1.0)
    def fx(r, x):
        r.update(np.array([[0.0, 0, 0], [x, 0, 0]]), np.array([1.0, 0.0]), np.array([1, 1], np.int8), tree=False)
        return r.forces(targets=np.array([1]))[0, 0]
    assert abs(fx(split, 5.0)) < 0.5 * abs(fx(full, 5.0))
    assert fx(split, 40.0) / fx(full, 40.0) > 0.8
    # pm_grid.rs:442-460: + attracted, - repelled by a + mass
    r = RefCpu(orc, 30.0, 32, 100.0, 0.5, 0.5, 0.0, 1.0, 1.0)
    r.update(np.array([[0.0, 0, 0], [10.0, 0, 0], [10.0, 0, 0]]), np.array([1.0, 0.0, 0.0]),
             np.array([1, 1, -1], np.int8), tree=False)
    a = r.forces(targets=np.array([1, 2]))
    assert a[0, 0] < 0 and a[1, 0] > 0


def test_ref_pm_isotropy(orc):
    # treepm_isotropy.rs:52-133: forces on a shell point radially within 5 degrees at 64^3
    n, R = 200, 30.0
    i = np.arange(n)
    y = 1 - i / (n - 1) * 2
    rr = np.sqrt(1 - y * y)
    th = np.pi * (3 - np.sqrt(5)) * i
    shell = np.stack([R * rr * np.cos(th), R * y, R * rr * np.sin(th)], 1)
    pos = np.concatenate([[[0.0, 0, 0]], shell])
    mass = np.concatenate([[100.0], np.zeros(n)])
    r = RefCpu(orc, 30.0, 64, 200.0, 0.5, 0.5, 0.0, 1.0, 1.0)
    r.update(pos, mass, np.ones(n + 1, np.int8), tree=False)
    a = r.forces(targets=np.arange(1, n + 1))
    radial = -shell / R
    cosang = (a * radial).sum(1) / np.linalg.norm(a, axis=1)
    ang = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
    assert ang.mean() < 5.0
