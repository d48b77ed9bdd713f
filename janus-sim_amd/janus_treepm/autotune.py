"""Autotuner for the walk's leaf size and group size and the FFT tile.

The tuner times candidate (leaf_size, group_size) pairs on a sub-box of the caller's particles and
keeps the fastest whose tree error matches the default's.  The sub-box is a cube of edge L / k cut
from the particle set and treated as its own periodic box with G / k cells, so the cell size, the
number density and the clustering match the full problem while a force evaluation costs 1 / k^3.
Sub-box timings do not always rank candidates as the full problem does: on the 4.1e7-particle,
512^3 box, leaf 8 beats leaf 4 by 16 % on 1/8- and 1/64-volume sub-boxes yet is 15 % slower on the
whole box (a smaller tree stays in the caches).  The default k = 1 therefore times the caller's own
particles (k = 1 keeps them all, shifted periodically); k > 1 is for problems too large to
evaluate a few times per candidate.

Tree error of a candidate = median relative difference between its accelerations and those of the
same context at theta = 0 (the direct short-range sum with the same mesh part): the mesh part is
common to every candidate, so this isolates what leaf and group size change (the opening test is
applied to the group box and the bucket size sets which sources stay particles).
"""
from __future__ import annotations

import time

import numpy as np

from . import FP32, ForceContext, DeviceArray, default_params, to_fixed

DEFAULT_CANDIDATES = ((4, 64), (2, 64), (8, 64), (4, 32), (8, 32), (2, 32), (4, 16), (8, 16))
FFT_TILES = (11, 12, 13)  # log2 points per FFT workgroup (jtp_params.fft_tile); all give the same bits


def sub_box(pos: np.ndarray, species: np.ndarray, box_size: float, n_cells: int, k: int = 1, seed: int = 0):
    """Particles of a random cube of edge L/k (positions in [-L/2, L/2)), re-centred as their own box."""
    rng = np.random.default_rng(seed)
    edge = box_size / k
    lo = rng.uniform(-box_size / 2, box_size / 2, 3)
    d = (pos - lo) % box_size  # offset from the cube's lower corner, periodic
    sel = np.all(d < edge, axis=1)
    return d[sel] - edge / 2, species[sel], edge, max(16, n_cells // k)


def _median_rel(a, b):
    return float(np.median(np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)))


def autotune(pos: np.ndarray, species: np.ndarray, params, m_plus: float, m_minus_abs: float, k: int = 1,
             candidates=DEFAULT_CANDIDATES, tolerance: float = 1.05, repeats: int = 3, seed: int = 0) -> dict:
    """Pick (leaf_size, group_size) for `params` (FP32 form) on a 1/k^3 sub-box of the particles.

    Returns {"leaf_size", "group_size", "fft_tile", "table": [(leaf, group, ms, tree_error), ...],
    "fft_table": [(tile, ms of a PM-only evaluation), ...], "sample": n}.
    A candidate qualifies if its tree error is <= tolerance x the error of params' own (leaf, group).
    """
    x, sp, L, G = sub_box(pos, species, params.box_size, params.n_cells, k, seed)
    if len(sp) < 1000:
        raise ValueError("sub-box holds too few particles to tune on")
    scale = L / params.box_size

    def make(leaf, group, theta=None, fft_tile=None, tree=None):
        p = default_params(L, G)
        for name in ("r_s", "r_cut", "softening"):
            setattr(p, name, getattr(params, name))
        p.theta = params.theta if theta is None else theta
        p.theta_cut = params.theta_cut if theta is None else theta
        p.g_constant, p.k_min = params.g_constant, params.k_min
        p.split_mode, p.gradient, p.deconvolve = params.split_mode, params.gradient, params.deconvolve
        p.interlace, p.pm_enabled, p.tree_enabled = params.interlace, params.pm_enabled, params.tree_enabled
        p.leaf_size, p.group_size, p.precision = leaf, group, FP32
        p.fft_tile = params.fft_tile if fft_tile is None else fft_tile
        if tree is not None:
            p.tree_enabled = tree
        if p.r_cut > L / 5:
            raise ValueError(f"sub-box edge {L:g} is too small for r_cut {p.r_cut:g}; use a smaller k")
        return ForceContext(p, m_plus, m_minus_abs)

    n = len(sp)
    dpos = DeviceArray.from_host(to_fixed(x, L))
    dsp = DeviceArray.from_host(sp.astype(np.int8))
    dacc = DeviceArray((n, 3), np.float32)

    ref_ctx = make(params.leaf_size, params.group_size, theta=0.0)
    ref_ctx.compute_forces_device(n, dpos, dsp, dacc)
    ref = dacc.download().astype(np.float64)
    ref_ctx.close()

    table = []
    for leaf, group in dict.fromkeys(tuple(c) for c in candidates + ((params.leaf_size, params.group_size),)):
        ctx = make(leaf, group)
        ctx.compute_forces_device(n, dpos, dsp, dacc)  # warm-up (allocations)
        best = np.inf
        for _ in range(repeats):
            t = time.perf_counter()
            ctx.compute_forces_device(n, dpos, dsp, dacc)
            best = min(best, time.perf_counter() - t)
        err = _median_rel(dacc.download().astype(np.float64), ref)
        table.append((leaf, group, best * 1e3, err))
        ctx.close()
    base = [row for row in table if row[0] == params.leaf_size and row[1] == params.group_size][0]
    ok = [row for row in table if row[3] <= tolerance * base[3]]
    pick = min(ok, key=lambda row: row[2])
    # FFT tile: PM-only evaluations (tree off) of the same particles, fastest wins (results identical)
    fft_table = []
    for tile in FFT_TILES:
        ctx = make(pick[0], pick[1], fft_tile=tile, tree=0)
        ctx.compute_forces_device(n, dpos, dsp, dacc)
        best = np.inf
        for _ in range(repeats):
            t = time.perf_counter()
            ctx.compute_forces_device(n, dpos, dsp, dacc)
            best = min(best, time.perf_counter() - t)
        fft_table.append((tile, best * 1e3))
        ctx.close()
    tile = min(fft_table, key=lambda row: row[1])[0]
    return {"leaf_size": pick[0], "group_size": pick[1], "fft_tile": tile, "table": table, "fft_table": fft_table,
            "sample": n, "sub_box": L, "volume_fraction": scale ** 3}
