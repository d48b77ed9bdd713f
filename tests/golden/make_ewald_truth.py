"""Writes FP64 Ewald truth fixtures for seeded target subsamples of the named
workloads (oracle/jtp_oracle.cpp orc_ewald: real space + structure factors,
Plummer-corrected nearest images).  The workloads are regenerated bit-identically
from their seeds; the fixture stores the SHA-256 of the fixed-point inputs so a
consumer can verify it is comparing against the same particles.

Run: python tests/golden/make_ewald_truth.py <name> [n_targets]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "janus-sim_amd"))
sys.path.insert(0, ROOT)

import janus_treepm as j  # noqa: E402
from janus_treepm import synth  # noqa: E402
from oracle import oracle_lib  # noqa: E402

WORKLOADS = {
    # the headline box: clustered N = 2.2e7, L = 250, PM 256^3
    "clustered_22M": lambda: synth.clustered(n_total=22_000_000, box_size=250.0, n_cells=256, n_halos=2000, seed=7),
    # the same recipe at 1/64 volume (same density, cell size, r_s, eps): quick accuracy studies
    "clustered_344k": lambda: synth.clustered(n_total=343_750, box_size=62.5, n_cells=64, n_halos=31, seed=99),
    # the small reference-CPU-sized box: lattice + Zel'dovich, 2 x 64^3, L = 250, PM 64^3
    "lattice_64": lambda: synth.lattice_zeldovich(n_side=64, box_size=250.0, n_cells=64, seed=42),
}


def make(name, n_targets=10_000, seed=2026):
    w = WORKLOADS[name]()
    if name == "clustered_344k":
        w.softening = synth.default_softening(250.0, 22_000_000)  # keep the full box's softening
    pos = j.to_fixed(w.pos, w.box_size)
    digest = w.digest(pos)
    rng = np.random.default_rng(seed)
    tg = np.sort(rng.choice(w.n, size=min(n_targets, w.n), replace=False)).astype(np.int64)
    orc = oracle_lib.load()
    rc = w.box_size / 4.0
    t0 = time.time()
    acc = orc.ewald(pos, w.species, w.m_plus, w.m_minus_abs, w.box_size, w.softening, targets=tg, rc=rc)
    dt = time.time() - t0
    out = os.path.join(HERE, f"ewald_{name}.npz")
    np.savez_compressed(out, targets=tg, acc=acc, digest=np.array(digest), softening=w.softening,
                        box_size=w.box_size, rc=rc, seconds=dt)
    print(f"{name}: {len(tg)} targets, N={w.n}, {dt:.1f} s -> {out}")


if __name__ == "__main__":
    make(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 10_000)
