"""ctypes loader for liboracle.so (ORACLE: test infrastructure only).

Wraps the C restatements in oracle/jtp_oracle.cpp with numpy-friendly helpers.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# JTP_ORACLE_LIB selects another build of the same source (the sanitizer build, oracle/Makefile `asan`)
LIB = os.environ.get("JTP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        vp, i64, i32, dbl, flt = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_float
        sig = {
            "orc_fixed_from_f64": (None, [i64, vp, dbl, vp]),
            "orc_morton_fixed": (None, [i64, vp, vp]),
            "orc_morton_f32_ref": (None, [i64, vp, flt, flt, vp]),
            "orc_morton30_f32_ref": (None, [i64, vp, flt, vp, vp]),
            "orc_radix_sort_ref": (None, [i64, vp, vp]),
            "orc_karras_ref": (None, [i64, vp, vp]),
            "orc_node_sums": (None, [i64, vp, vp, vp, vp, vp]),
            "orc_deposit": (None, [i64, vp, vp, dbl, dbl, i32, vp]),
            "orc_fft3d": (None, [i32, vp, vp, i32]),
            "orc_poisson": (None, [i32, dbl, dbl, dbl, i32, i32, i32, vp, vp]),
            "orc_interp": (None, [i64, vp, vp, i32, dbl, vp, i32, vp]),
            "orc_short_direct": (None, [i64, vp, vp, dbl, dbl, dbl, dbl, dbl, dbl, i32, i64, vp, vp]),
            "orc_ewald": (None, [i64, vp, vp, dbl, dbl, dbl, dbl, dbl, dbl, i32, i64, vp, vp]),
            "orc_ewald_sc_potential": (dbl, [dbl, dbl, i32, i32]),
            "orc_refcpu_create": (vp, [dbl, i32, dbl, dbl, dbl, dbl, dbl, dbl]),
            "orc_refcpu_destroy": (None, [vp]),
            "orc_refcpu_update": (None, [vp, i64, vp, vp, vp, i32, i32]),
            "orc_refcpu_forces": (None, [vp, i64, vp, i32, i32, vp]),
            "orc_num_threads": (C.c_int, []),
            "orc_segregation": (None, [i64, vp, vp, i32, vp]),
            "orc_splitmix64": (None, [C.c_uint64, i64, vp]),
            "orc_ic": (i64, [i32, dbl, dbl, C.c_uint64, dbl, dbl, dbl, vp, vp, vp]),
            "orc_segregation_f32_ref": (None, [i64, vp, vp, dbl, i32, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a

    # helpers -------------------------------------------------------------
    @staticmethod
    def _p(a):
        return a.ctypes.data if a is not None else None

    def fixed(self, x, L):
        x = np.ascontiguousarray(x, np.float64)
        out = np.empty(x.shape, np.uint32)
        self.lib.orc_fixed_from_f64(x.size, x.ctypes.data, L, out.ctypes.data)
        return out

    def morton(self, pos_fixed):
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        k = np.empty(len(p), np.uint64)
        self.lib.orc_morton_fixed(len(p), p.ctypes.data, k.ctypes.data)
        return k

    def morton_f32_ref(self, pos_f32, box_half, inv_cell):
        p = np.ascontiguousarray(pos_f32, np.float32).reshape(-1, 3)
        k = np.empty(len(p), np.uint64)
        self.lib.orc_morton_f32_ref(len(p), p.ctypes.data, box_half, inv_cell, k.ctypes.data)
        return k

    def morton30_f32_ref(self, pos_f32, box_half):
        """(keys, indices) of compute_morton_f32: 30-bit key << 32 | index."""
        p = np.ascontiguousarray(pos_f32, np.float32).reshape(-1, 3)
        k = np.empty(len(p), np.uint64)
        idx = np.empty(len(p), np.int32)
        self.lib.orc_morton30_f32_ref(len(p), p.ctypes.data, box_half, k.ctypes.data, idx.ctypes.data)
        return k, idx

    def radix_sort(self, keys):
        k = np.ascontiguousarray(keys, np.uint64)
        perm = np.empty(len(k), np.int32)
        self.lib.orc_radix_sort_ref(len(k), k.ctypes.data, perm.ctypes.data)
        return perm

    def karras(self, sorted_keys):
        k = np.ascontiguousarray(sorted_keys, np.uint64)
        n = len(k)
        out = np.empty((max(n - 1, 0), 4), np.int32)
        if n >= 2:
            self.lib.orc_karras_ref(n, k.ctypes.data, out.ctypes.data)
        return out

    def node_sums(self, spos, ssp, nodes):
        spos = np.ascontiguousarray(spos, np.uint32).reshape(-1, 3)
        ssp = np.ascontiguousarray(ssp, np.int8)
        nodes = np.ascontiguousarray(nodes, np.int32)
        n = len(ssp)
        sums = np.empty((n - 1, 8), np.uint64)
        boxes = np.empty((n - 1, 6), np.uint32)
        self.lib.orc_node_sums(n, spos.ctypes.data, ssp.ctypes.data, nodes.ctypes.data, sums.ctypes.data,
                               boxes.ctypes.data)
        return sums, boxes

    def splitmix64(self, seed, count):
        out = np.zeros(count, np.uint64)
        self.lib.orc_splitmix64(seed, count, out.ctypes.data)
        return out

    def ic(self, n, L, z_init=10.0, seed=42, mu=19.0, n_s=0.965, delta_rms=0.10):
        """(pos [n^3, 3] in [-L/2, L/2], vel, sign, n_plus) of the Zel'dovich lattice ICs."""
        pos = np.zeros((n ** 3, 3))
        vel = np.zeros((n ** 3, 3))
        sign = np.zeros(n ** 3, np.int8)
        n_plus = self.lib.orc_ic(n, L, z_init, seed, mu, n_s, delta_rms, pos.ctypes.data, vel.ctypes.data,
                                 sign.ctypes.data)
        return pos, vel, sign, int(n_plus)

    def segregation(self, pos_fixed, sp, n_cells):
        pos = np.ascontiguousarray(pos_fixed, np.uint32)
        sp = np.ascontiguousarray(sp, np.int8)
        out = np.zeros(3)
        self.lib.orc_segregation(len(sp), pos.ctypes.data, sp.ctypes.data, n_cells, out.ctypes.data)
        return out

    def segregation_f32_ref(self, pos_f32, sp, L, n_cells):
        pos = np.ascontiguousarray(pos_f32, np.float32)
        sp = np.ascontiguousarray(sp, np.int8)
        out = np.zeros(3)
        self.lib.orc_segregation_f32_ref(len(sp), pos.ctypes.data, sp.ctypes.data, L, n_cells, out.ctypes.data)
        return out

    def deposit(self, pos_fixed, sp, mp, mm, G):
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        s = np.ascontiguousarray(sp, np.int8)
        rho = np.empty(G ** 3, np.float64)
        self.lib.orc_deposit(len(s), p.ctypes.data, s.ctypes.data, mp, mm, G, rho.ctypes.data)
        return rho

    def fft3d(self, z, inverse=False):
        z = np.asarray(z, np.complex128)
        G = z.shape[0]
        re = np.ascontiguousarray(z.real)
        im = np.ascontiguousarray(z.imag)
        # oracle layout: index i + G(j + G k) with i along the LAST numpy axis for [k][j][i] arrays
        self.lib.orc_fft3d(G, re.ctypes.data, im.ctypes.data, 1 if inverse else 0)
        return re + 1j * im

    def poisson(self, rho, G, L, g, r_s, k_min=0, deconv=0, match=-1):
        """match: -1 none, 0 / 1 = Green function matched to the 4- / 2-point gradient stencil."""
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        phi = np.empty_like(r)
        self.lib.orc_poisson(G, L, g, r_s, k_min, deconv, match, r.ctypes.data, phi.ctypes.data)
        return phi

    def interp(self, pos_fixed, sp, G, L, phi, four=True):
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        s = np.ascontiguousarray(sp, np.int8)
        ph = np.ascontiguousarray(phi, np.float64).reshape(-1)
        acc = np.empty((len(s), 3), np.float64)
        self.lib.orc_interp(len(s), p.ctypes.data, s.ctypes.data, G, L, ph.ctypes.data, 1 if four else 0,
                            acc.ctypes.data)
        return acc

    def short_direct(self, pos_fixed, sp, qp, qm, L, eps, rcut, rs, mode=0, targets=None):
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        s = np.ascontiguousarray(sp, np.int8)
        t = None if targets is None else np.ascontiguousarray(targets, np.int64)
        nt = len(s) if t is None else len(t)
        acc = np.empty((nt, 3), np.float64)
        self.lib.orc_short_direct(len(s), p.ctypes.data, s.ctypes.data, qp, qm, L, eps, rcut, rs, mode, nt,
                                  self._p(t), acc.ctypes.data)
        return acc

    def ewald(self, pos_fixed, sp, qp, qm, L, eps, targets=None, rc=None, tol_arg=4.6):
        """qp / qm: G |m| of species +1 / -1 (both >= 0)."""
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        s = np.ascontiguousarray(sp, np.int8)
        t = None if targets is None else np.ascontiguousarray(targets, np.int64)
        nt = len(s) if t is None else len(t)
        if rc is None:
            rc = L / 4.0
        alpha = tol_arg / rc
        nk = int(np.ceil(2.0 * alpha * tol_arg * L / (2.0 * np.pi)))
        acc = np.empty((nt, 3), np.float64)
        self.lib.orc_ewald(len(s), p.ctypes.data, s.ctypes.data, qp, qm, L, eps, alpha, rc, nk, nt, self._p(t),
                           acc.ctypes.data)
        return acc

    def ewald_custom(self, pos_fixed, sp, qp, qm, L, eps, alpha, rc, nk, targets=None):
        p = np.ascontiguousarray(pos_fixed, np.uint32).reshape(-1, 3)
        s = np.ascontiguousarray(sp, np.int8)
        t = None if targets is None else np.ascontiguousarray(targets, np.int64)
        nt = len(s) if t is None else len(t)
        acc = np.empty((nt, 3), np.float64)
        self.lib.orc_ewald(len(s), p.ctypes.data, s.ctypes.data, qp, qm, L, eps, alpha, rc, nk, nt, self._p(t),
                           acc.ctypes.data)
        return acc


class RefCpu:
    """Restatement of the reference CPU TreePMForce (treepm_force.rs:17-146)."""

    def __init__(self, orc: Oracle, r_cut, G, L, theta, eps, r_s, g_pm, g_tree):
        self.orc = orc
        self.h = orc.lib.orc_refcpu_create(r_cut, G, L, theta, eps, r_s, g_pm, g_tree)

    def update(self, pos, mass, sign, pm=True, tree=True):
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.mass = np.ascontiguousarray(mass, np.float64)
        self.sign = np.ascontiguousarray(sign, np.int8)
        self.pm, self.tree = int(pm), int(tree)
        self.orc.lib.orc_refcpu_update(self.h, len(self.sign), self.pos.ctypes.data, self.mass.ctypes.data,
                                       self.sign.ctypes.data, self.pm, self.tree)

    def forces(self, targets=None):
        t = None if targets is None else np.ascontiguousarray(targets, np.int64)
        nt = len(self.sign) if t is None else len(t)
        acc = np.empty((nt, 3), np.float64)
        self.orc.lib.orc_refcpu_forces(self.h, nt, None if t is None else t.ctypes.data, self.pm, self.tree,
                                       acc.ctypes.data)
        return acc

    def __del__(self):
        try:
            self.orc.lib.orc_refcpu_destroy(self.h)
        except Exception:
            pass


_oracle = None


def load() -> Oracle:
    global _oracle
    if _oracle is None:
        if not os.path.exists(LIB):
            build()
        _oracle = Oracle(C.CDLL(LIB))
    return _oracle
