"""Seeded synthetic two-species particle sets for the TreePM workloads.

Each generator returns a ``Workload`` with host positions in [-L/2, L/2)
(float64, shape [n, 3]), species (+1 / -1, int8) and the force parameters.
Only IEEE-exact numpy operations (+ - * / sqrt, sorting, bincount, FFT) touch
per-particle data; transcendental functions are evaluated with the scalar
``math`` module on small tables, so a seed gives bit-identical particles on any
host of this image (numpy's SIMD transcendental ufuncs may differ in the last bit
between CPU models).  The softening default follows config.rs:191-196 (L / N^{1/3} / 30 with N the
total particle count); the split scale is r_s = 1.25 cells with r_cut = 4.5 r_s.

* ``lattice_zeldovich``   - lattice + Zel'dovich displacement from a Gaussian field with a
                            power-law displacement spectrum; m- carries the negated field on
                            the lattice offset by half a cell (anti-correlated species).
* ``clustered``           - NFW halos (60 % of m+) + uniform m+, with m- rejection-sampled
                            where the Gaussian-smoothed m+ density is below its median.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    pos: np.ndarray          # float64 [n, 3] in [-L/2, L/2)
    species: np.ndarray      # int8 [n]
    box_size: float
    n_cells: int
    r_s: float
    r_cut: float
    theta: float
    softening: float
    m_plus: float
    m_minus_abs: float
    leaf_size: int = 4
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.species)

    def digest(self, pos_fixed: np.ndarray) -> str:
        """SHA-256 of the fixed-point positions and species (identifies the exact inputs)."""
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(pos_fixed, np.uint32).tobytes())
        h.update(np.ascontiguousarray(self.species, np.int8).tobytes())
        return h.hexdigest()


def default_softening(box_size: float, n_total: int) -> float:
    return box_size / n_total ** (1.0 / 3.0) / 30.0


def _wrap(x, L):
    return x - L * np.floor(x / L + 0.5)


def zeldovich_displacement(n: int, box_size: float, seed: int, index: float = -2.5, rms_cells: float = 0.3):
    """Displacement field psi on an n^3 lattice: psi_k = i k / k^2 delta_k with |delta_k|^2 ~ P(k) k^2 such
    that the displacement spectrum is P(k) ~ k^index; scaled to rms |psi| = rms_cells * (L / n)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    white = rng.standard_normal((n, n, n))
    wk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    # sqrt(P_psi(k)) with P_psi ~ k^index, from a table over the integer values of k^2
    k2i = np.rint(k2).astype(np.int64)
    table = np.array([0.0] + [math.pow(float(v), index / 4.0) for v in range(1, int(k2i.max()) + 1)])
    amp = table[k2i]
    amp[0, 0, 0] = 0.0
    kmag = np.sqrt(k2)
    psi = []
    for K in (KX, KY, KZ):
        comp = np.fft.irfftn(1j * (K / kmag) * amp * wk, s=(n, n, n), axes=(0, 1, 2))
        psi.append(comp)
    psi = np.stack(psi, axis=-1)
    rms = np.sqrt(np.mean(np.sum(psi ** 2, axis=-1)))
    psi *= rms_cells * (box_size / n) / rms
    return psi


def lattice_zeldovich(n_side: int = 64, box_size: float = 250.0, n_cells: int = 64, seed: int = 42,
                      m_plus: float = 1.0, m_minus_abs: float = 19.0, theta: float = 0.5,
                      leaf_size: int = 4, name: str = "lattice_zeldovich") -> Workload:
    h = box_size / n_side
    psi = zeldovich_displacement(n_side, box_size, seed)
    g = (np.arange(n_side) + 0.5) * h - box_size / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    q = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    p_plus = _wrap(q + psi.reshape(-1, 3), box_size)
    p_minus = _wrap(q + 0.5 * h - psi.reshape(-1, 3), box_size)
    pos = np.concatenate([p_plus, p_minus])
    sp = np.concatenate([np.ones(len(p_plus), np.int8), -np.ones(len(p_minus), np.int8)])
    cell = box_size / n_cells
    r_s = 1.25 * cell
    return Workload(name, pos, sp, box_size, n_cells, r_s, 4.5 * r_s, theta,
                    default_softening(box_size, len(sp)), m_plus, m_minus_abs, leaf_size)


def _nfw_inverse_cdf(c: float, u: np.ndarray) -> np.ndarray:
    """Radius / r_200 with enclosed-mass fraction u for an NFW profile of concentration c."""
    x = np.linspace(0.0, 1.0, 2049)
    mu = np.array([math.log1p(c * t) - c * t / (1.0 + c * t) for t in x])
    cdf = mu / mu[-1]
    return np.interp(u, cdf, x)


def _cic_density(pos, box_size, G):
    u = (pos / box_size + 0.5) * G
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    rho = np.zeros(G ** 3)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                idx = ((i0[:, 0] + dx) % G) + G * (((i0[:, 1] + dy) % G) + G * ((i0[:, 2] + dz) % G))
                rho += np.bincount(idx, weights=wx * wy * wz, minlength=G ** 3)
    return rho.reshape(G, G, G)  # [z][y][x]


def clustered(n_total: int = 22_000_000, box_size: float = 250.0, n_cells: int = 256, n_halos: int = 2000,
              seed: int = 7, m_plus: float = 1.0, m_minus_abs: float = 19.0, theta: float = 0.5,
              halo_fraction: float = 0.6, leaf_size: int = 4, name: str = "clustered",
              n_plus: int | None = None) -> Workload:
    """Clustered two-species box: 60 % of m+ in NFW halos (dN/dM ~ M^-1.9 on [1e-5, 1e-2] of the m+
    mass, concentration U[5, 15], r_200 at 200 x the mean m+ density), the rest uniform; m- rejection
    sampled where the CIC density of m+ smoothed by a 2-cell Gaussian is below its median.  n_plus
    defaults to half of n_total."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = box_size
    if n_plus is None:
        n_plus = n_total // 2
    n_minus = n_total - n_plus
    # halo masses ~ M^-1.9 on [1e-5, 1e-2] of the total m+ mass (inverse CDF of a power law)
    a, lo, hi = -1.9, 1e-5, 1e-2
    u = rng.random(n_halos)
    e = a + 1.0
    mh = np.array([math.pow(math.pow(lo, e) + v * (math.pow(hi, e) - math.pow(lo, e)), 1.0 / e) for v in u])
    n_in_halos = int(round(halo_fraction * n_plus))
    counts = np.floor(n_in_halos * mh / mh.sum()).astype(np.int64)
    counts[: n_in_halos - counts.sum()] += 1
    conc = rng.uniform(5.0, 15.0, n_halos)
    centres = rng.uniform(-L / 2, L / 2, (n_halos, 3))
    rho_bar = n_plus / L ** 3
    r200 = np.array([math.pow(3.0 * float(c) / (800.0 * math.pi * rho_bar), 1.0 / 3.0) for c in counts])
    parts = []
    for h in range(n_halos):
        k = int(counts[h])
        if k == 0:
            continue
        r = r200[h] * _nfw_inverse_cdf(conc[h], rng.random(k))
        v = rng.standard_normal((k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        parts.append(centres[h] + r[:, None] * v)
    halo_pos = np.concatenate(parts) if parts else np.zeros((0, 3))
    uni = rng.uniform(-L / 2, L / 2, (n_plus - len(halo_pos), 3))
    p_plus = _wrap(np.concatenate([halo_pos, uni]), L)
    # m-: rejection sampling where the smoothed m+ density is below its median
    G = n_cells
    rho = _cic_density(p_plus, L, G)
    sig = 2.0  # cells; separable Gaussian filter exp(-sig^2 k^2 / 2)
    k = np.fft.fftfreq(G) * 2 * np.pi
    gk = np.array([math.exp(-0.5 * sig ** 2 * float(v) ** 2) for v in k])
    gz = gk[: G // 2 + 1].copy()
    gz[G // 2] = gk[G // 2]
    filt = gk[:, None, None] * gk[None, :, None] * gz[None, None, :]
    sm = np.fft.irfftn(np.fft.rfftn(rho) * filt, s=(G, G, G), axes=(0, 1, 2))
    med = np.median(sm)
    low = sm < med
    out = []
    need = n_minus
    while need > 0:
        cand = rng.uniform(-L / 2, L / 2, (int(need * 2.2) + 1024, 3))
        ci = np.floor((cand / L + 0.5) * G).astype(np.int64) % G
        ok = low[ci[:, 2], ci[:, 1], ci[:, 0]]
        acc = cand[ok][:need]
        out.append(acc)
        need -= len(acc)
    p_minus = np.concatenate(out)
    pos = np.concatenate([p_plus, p_minus])
    sp = np.concatenate([np.ones(n_plus, np.int8), -np.ones(n_minus, np.int8)])
    cell = L / n_cells
    r_s = 1.25 * cell
    return Workload(name, pos, sp, L, n_cells, r_s, 4.5 * r_s, theta, default_softening(L, n_total),
                    m_plus, m_minus_abs, leaf_size, extra={"n_halos": n_halos, "median_smoothed": float(med)})


def ensemble_c3(n_boxes: int = 48, seed: int = 1000) -> list:
    """Uneven ensemble of lattice + Zel'dovich two-species boxes: N_b = 2 n^3 with n = round(2^u),
    u ~ U[5, 7] (N log2-uniform in [2*32^3, 2*128^3]); G = 2^round(log2 N^(1/3)) in {32, 64, 128};
    L ~ U[100, 500]; |m-| / m+ ~ U[5, 40] with m+ = 1; box b uses the lattice recipe with seed + b.
    The per-box draws come from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for b in range(n_boxes):
        u = rng.uniform(5.0, 7.0)
        L = float(rng.uniform(100.0, 500.0))
        ratio = float(rng.uniform(5.0, 40.0))
        n = int(round(2.0 ** u))
        N = 2 * n ** 3
        G = int(2 ** int(round(np.log2(N ** (1.0 / 3.0)))))
        w = lattice_zeldovich(n_side=n, box_size=L, n_cells=G, seed=seed + b, m_plus=1.0, m_minus_abs=ratio,
                              name=f"c3_box{b}")
        out.append(w)
    return out

