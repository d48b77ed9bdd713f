"""Python binding of libjanus_treepm (the MI355X TreePM force step).

The library is a C ABI (include/janus_treepm.h); this module loads it with
ctypes and mirrors the reference's Rust-side interface for this path:

* ``TreePMForce``      - treepm_force.rs:17-146 (new / update / compute_all_forces),
                         particle arrays in host memory, FP64 results.
* ``PoissonSolver``    - cufft_ffi.rs:71-354 (CuFFTPoisson / solve_device), device grids.
* ``ForceContext``     - the device-pointer force call of GpuNBodyTwoPass
                         (nbody_gpu_twopass.rs:3961-4138) and its integrator step
                         step_treepm_gpu (:4486-4649) as ``step_device``.

Errors from the library raise ``JtpError`` (the reference maps a -1 status to
``Err(String)``, cufft_ffi.rs:97-99).  There is no CPU fallback: if the shared
library is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libjanus_treepm.so")

JTP_OK, JTP_ERR_ARG, JTP_ERR_HIP, JTP_ERR_ALLOC, JTP_ERR_OVERFLOW, JTP_ERR_STATE = 0, -1, -2, -3, -4, -5
FP32, FP64 = 0, 1
SPLIT_ERFC_GAUSS, SPLIT_ERFC, SPLIT_POLY4 = 0, 1, 2
GRAD_4PT, GRAD_2PT = 0, 1
STAGE_KEYS, STAGE_PERM, STAGE_NODES, STAGE_MONO, STAGE_RHO, STAGE_PHI, STAGE_ACC_PM, STAGE_SUMS, STAGE_BOXES = (
    1, 2, 3, 4, 5, 6, 7, 8, 9)
STAGE_RHO2, STAGE_PHI2 = 10, 11
STEP_DKD, STEP_KDK, STEP_GRAPH = 0, 1, 0x100
ENS_SEQUENTIAL, ENS_WALK_PER_BOX, ENS_GRAPH = 0x1, 0x2, 0x100

# every symbol declared in include/janus_treepm.h
EXPORTED = [
    "jtp_default_params", "jtp_create", "jtp_destroy", "jtp_last_error", "jtp_version", "jtp_set_species_mass",
    "jtp_compute_forces", "jtp_pos_to_fixed_f64", "jtp_pos_to_fixed_f32", "jtp_pm_solve", "jtp_morton_keys_f32",
    "jtp_get_stats", "jtp_stage_size", "jtp_stage_copy", "jtp_alloc", "jtp_free", "jtp_copy_h2d",
    "jtp_copy_d2h", "jtp_copy_d2d", "jtp_device_synchronize", "jtp_device_count", "jtp_step",
    "jtp_ensemble_create", "jtp_ensemble_destroy", "jtp_ensemble_last_error", "jtp_ensemble_compute_forces",
    "jtp_ensemble_get_stats", "jtp_diagnostics", "jtp_potential_energy", "jtp_device_memory",
    "jtp_segregation", "jtp_snap_meta_default", "jtp_snapshot_write", "jtp_snapshot_read_meta", "jtp_snapshot_read",
    "jtp_io_last_error", "jtp_ic_default_params", "jtp_generate_ics", "jtp_params_from_config",
    "jtp_morton30_keys_f32", "jtp_pm_solve_noalloc",
]


class JtpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"janus_treepm error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("box_size", C.c_double), ("n_cells", C.c_int32), ("k_min", C.c_int32),
        ("r_s", C.c_double), ("r_cut", C.c_double), ("theta", C.c_double), ("softening", C.c_double),
        ("g_constant", C.c_double), ("leaf_size", C.c_int32), ("group_size", C.c_int32),
        ("precision", C.c_int32), ("split_mode", C.c_int32), ("gradient", C.c_int32),
        ("deconvolve", C.c_int32), ("pm_enabled", C.c_int32), ("tree_enabled", C.c_int32),
        ("interlace", C.c_int32), ("theta_cut", C.c_double), ("fft_tile", C.c_int32),
    ]


class ConfigFields(C.Structure):
    """jtp_config_fields: the YAML fields of JanusConfig that set the force (config.rs:21-40,94-102) plus the
    optional pm_grid.split_scale_cells / pm_grid.r_cut_over_r_s (absent: <= 0)."""
    _fields_ = [
        ("box_size_mpc", C.c_double), ("n_particles", C.c_int64), ("theta", C.c_double),
        ("softening_mpc", C.c_double), ("n_cells", C.c_int32), ("k_min", C.c_int32),
        ("split_scale_cells", C.c_double), ("r_cut_over_r_s", C.c_double),
    ]


class Diag(C.Structure):
    _fields_ = [("kinetic", C.c_double * 2), ("momentum", (C.c_double * 3) * 2), ("pm_energy", C.c_double)]


SNAP_PRODUCTION, SNAP_V3 = 1, 3


class SnapMeta(C.Structure):
    """jtp_snap_meta (include/janus_treepm.h)."""
    _fields_ = [
        ("format", C.c_int32), ("n", C.c_int64), ("box_size", C.c_double), ("step", C.c_uint32), ("z", C.c_double),
        ("velocity_scale", C.c_double), ("a", C.c_double), ("t_gyr", C.c_double), ("h0", C.c_double),
        ("mu", C.c_double), ("omega_b", C.c_double), ("m_part_plus_base", C.c_double),
        ("m_part_minus_base", C.c_double), ("eps_plus_base", C.c_double), ("eps_minus_base", C.c_double),
        ("n_split_max", C.c_uint32), ("seed_ic", C.c_uint32), ("z_init", C.c_double), ("n_stars", C.c_uint64),
        ("z_start_run", C.c_double), ("sfr", C.c_double), ("rho_max", C.c_double), ("run_label", C.c_char * 256),
        ("mass_plus", C.c_float), ("mass_minus", C.c_float), ("eps_plus", C.c_float), ("eps_minus", C.c_float),
    ]


class IcParams(C.Structure):
    """jtp_ic_params (include/janus_treepm.h)."""
    _fields_ = [("box_size", C.c_double), ("n_grid", C.c_int32), ("z_init", C.c_double), ("seed", C.c_uint64),
                ("mu", C.c_double), ("n_s", C.c_double), ("delta_rms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_sort", C.c_double), ("ms_build", C.c_double), ("ms_deposit", C.c_double),
        ("ms_fft", C.c_double), ("ms_interp", C.c_double), ("ms_walk", C.c_double),
        ("interactions", C.c_uint64), ("interactions_mono", C.c_uint64), ("n_groups", C.c_int64),
        ("n_particles", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libjanus_treepm.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not built: run `make -C janus-sim_amd` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, i64, i32, sz, dbl = C.c_void_p, C.c_int64, C.c_int32, C.c_size_t, C.c_double
    sig = {
        "jtp_default_params": (None, [C.POINTER(Params), dbl, i32]),
        "jtp_create": (C.c_int, [C.POINTER(Params), C.POINTER(vp)]),
        "jtp_destroy": (C.c_int, [vp]),
        "jtp_last_error": (C.c_char_p, [vp]),
        "jtp_version": (C.c_char_p, []),
        "jtp_set_species_mass": (C.c_int, [vp, dbl, dbl]),
        "jtp_compute_forces": (C.c_int, [vp, i64, vp, vp, vp, vp]),
        "jtp_pos_to_fixed_f64": (C.c_int, [vp, i64, vp, vp, vp]),
        "jtp_pos_to_fixed_f32": (C.c_int, [vp, i64, vp, vp, vp]),
        "jtp_pm_solve": (C.c_int, [vp, vp, vp, vp]),
        "jtp_pm_solve_noalloc": (C.c_int, [C.POINTER(Params), vp, vp, vp]),
        "jtp_morton_keys_f32": (C.c_int, [i64, vp, C.c_float, C.c_float, vp, vp]),
        "jtp_morton30_keys_f32": (C.c_int, [i64, vp, C.c_float, C.c_float, vp, vp, vp]),
        "jtp_params_from_config": (C.c_int, [C.POINTER(ConfigFields), C.POINTER(Params)]),
        "jtp_get_stats": (C.c_int, [vp, C.POINTER(Stats)]),
        "jtp_stage_size": (C.c_int, [vp, i32, C.POINTER(sz)]),
        "jtp_stage_copy": (C.c_int, [vp, i32, vp, sz]),
        "jtp_alloc": (vp, [sz]),
        "jtp_free": (None, [vp]),
        "jtp_copy_h2d": (C.c_int, [vp, vp, sz]),
        "jtp_copy_d2h": (C.c_int, [vp, vp, sz]),
        "jtp_copy_d2d": (C.c_int, [vp, vp, sz]),
        "jtp_device_synchronize": (C.c_int, []),
        "jtp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "jtp_step": (C.c_int, [vp, i64, vp, vp, vp, vp, i32, dbl, dbl, dbl, dbl, vp]),
        "jtp_ensemble_create": (C.c_int, [i32, C.POINTER(Params), C.POINTER(dbl), C.POINTER(dbl), i32,
                                          C.POINTER(vp)]),
        "jtp_ensemble_destroy": (C.c_int, [vp]),
        "jtp_ensemble_last_error": (C.c_char_p, [vp]),
        "jtp_ensemble_compute_forces": (C.c_int, [vp, C.POINTER(i64), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                                  i32, vp]),
        "jtp_ensemble_get_stats": (C.c_int, [vp, i32, C.POINTER(Stats)]),
        "jtp_diagnostics": (C.c_int, [vp, i64, vp, vp, C.POINTER(Diag), vp]),
        "jtp_potential_energy": (C.c_int, [vp, i64, vp, vp, C.POINTER(dbl), vp]),
        "jtp_device_memory": (C.c_int, [C.POINTER(sz), C.POINTER(sz)]),
        "jtp_segregation": (C.c_int, [vp, i64, vp, vp, i32, C.POINTER(dbl), vp]),
        "jtp_snap_meta_default": (None, [C.POINTER(SnapMeta), i32, dbl]),
        "jtp_snapshot_write": (C.c_int, [C.c_char_p, C.POINTER(SnapMeta), vp, vp, i32, vp, vp]),
        "jtp_snapshot_read_meta": (C.c_int, [C.c_char_p, i32, C.POINTER(SnapMeta)]),
        "jtp_snapshot_read": (C.c_int, [C.c_char_p, i32, i64, vp, vp, i32, vp, dbl, vp]),
        "jtp_io_last_error": (C.c_char_p, []),
        "jtp_ic_default_params": (None, [C.POINTER(IcParams)]),
        "jtp_generate_ics": (C.c_int, [vp, C.POINTER(IcParams), vp, vp, i32, vp, C.POINTER(i64), vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def device_memory() -> tuple:
    """(free, total) bytes of the current device."""
    lib = load_library()
    f, t = C.c_size_t(0), C.c_size_t(0)
    rc = lib.jtp_device_memory(C.byref(f), C.byref(t))
    if rc != JTP_OK:
        raise JtpError(rc, "hipMemGetInfo failed")
    return f.value, t.value


def snap_meta(fmt: int, box_size: float, **fields) -> SnapMeta:
    """Metadata with the library's defaults (SnapshotHeaderV3::default), then `fields`."""
    m = SnapMeta()
    load_library().jtp_snap_meta_default(C.byref(m), int(fmt), float(box_size))
    for k, v in fields.items():
        setattr(m, k, v.encode() if k == "run_label" and isinstance(v, str) else v)
    return m


def ic_params(**fields) -> IcParams:
    """IC parameters with IcParams::default (ic_gen.rs:24-36), then `fields`."""
    p = IcParams()
    load_library().jtp_ic_default_params(C.byref(p))
    for k, v in fields.items():
        setattr(p, k, v)
    return p


def _io_check(rc):
    if rc != JTP_OK:
        raise JtpError(rc, load_library().jtp_io_last_error().decode())


def write_snapshot(path: str, meta: SnapMeta, d_pos_fixed, d_vel, d_species, vel_f64: bool = False, stream=None):
    """Write meta.n particles from device arrays (jtp_snapshot_write)."""
    ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
    _io_check(load_library().jtp_snapshot_write(path.encode(), C.byref(meta), ptr(d_pos_fixed), ptr(d_vel),
                                                int(vel_f64), ptr(d_species), stream))


def read_snapshot_meta(path: str, fmt: int) -> SnapMeta:
    m = SnapMeta()
    _io_check(load_library().jtp_snapshot_read_meta(path.encode(), int(fmt), C.byref(m)))
    return m


def read_snapshot(path: str, fmt: int, n: int, d_pos_fixed, d_vel, d_species, vel_f64: bool = False,
                  velocity_scale: float = 1.0, stream=None):
    """Load the file's n particles into device arrays (jtp_snapshot_read)."""
    ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
    _io_check(load_library().jtp_snapshot_read(path.encode(), int(fmt), int(n), ptr(d_pos_fixed), ptr(d_vel),
                                               int(vel_f64), ptr(d_species), float(velocity_scale), stream))


def device_count() -> int:
    lib = load_library()
    n = C.c_int(0)
    rc = lib.jtp_device_count(C.byref(n))
    return n.value if rc == JTP_OK else 0


class DeviceArray:
    """A device allocation holding a numpy-described array (jtp_alloc / jtp_free)."""

    def __init__(self, shape, dtype):
        self.lib = load_library()
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = self.lib.jtp_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise JtpError(JTP_ERR_ALLOC, f"jtp_alloc({self.nbytes}) failed")

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        if self.nbytes and self.lib.jtp_copy_h2d(self.ptr, a.ctypes.data, self.nbytes) != JTP_OK:
            raise JtpError(JTP_ERR_HIP, "copy to device failed")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes and self.lib.jtp_copy_d2h(out.ctypes.data, self.ptr, self.nbytes) != JTP_OK:
            raise JtpError(JTP_ERR_HIP, "copy to host failed")
        return out

    def free(self):
        if self.ptr:
            self.lib.jtp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_params(box_size: float, n_cells: int, **kw) -> Params:
    lib = load_library()
    p = Params()
    lib.jtp_default_params(C.byref(p), float(box_size), int(n_cells))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def params_from_config(box_size_mpc: float, n_cells: int, n_particles: int = 0, theta: float | None = None,
                       softening_mpc: float | None = None, k_min: int = 0, split_scale_cells: float | None = None,
                       r_cut_over_r_s: float | None = None) -> Params:
    """jtp_params from the YAML fields (jtp_params_from_config); None means the field is absent."""
    f = ConfigFields(float(box_size_mpc), int(n_particles), -1.0 if theta is None else float(theta),
                     0.0 if softening_mpc is None else float(softening_mpc), int(n_cells), int(k_min),
                     0.0 if split_scale_cells is None else float(split_scale_cells),
                     0.0 if r_cut_over_r_s is None else float(r_cut_over_r_s))
    p = Params()
    rc = load_library().jtp_params_from_config(C.byref(f), C.byref(p))
    if rc != JTP_OK:
        raise JtpError(rc, "invalid config fields")
    return p


class ForceContext:
    """One jtp context: the device force call on fixed-point particle arrays."""

    def __init__(self, params: Params, m_plus: float = 1.0, m_minus_abs: float = 1.0):
        self.lib = load_library()
        self.params = params
        h = C.c_void_p()
        rc = self.lib.jtp_create(C.byref(params), C.byref(h))
        if rc != JTP_OK:
            raise JtpError(rc, "jtp_create failed (invalid parameters or device error)")
        self.h = h
        self.set_species_mass(m_plus, m_minus_abs)

    @property
    def real(self):
        return np.float64 if self.params.precision == FP64 else np.float32

    def _check(self, rc):
        if rc != JTP_OK:
            raise JtpError(rc, self.lib.jtp_last_error(self.h).decode())

    def set_species_mass(self, m_plus: float, m_minus_abs: float):
        self._check(self.lib.jtp_set_species_mass(self.h, float(m_plus), float(m_minus_abs)))

    def _arg(self, d, dtype, count, name):
        """Raw pointer of a device argument; a DeviceArray must hold `count` elements of `dtype` (a wrong
        precision or a short buffer would make the kernels write past the allocation)."""
        if not isinstance(d, DeviceArray):
            return d
        if d.dtype != np.dtype(dtype):
            raise ValueError(f"{name}: dtype {d.dtype}, expected {np.dtype(dtype)}")
        if d.nbytes < count * np.dtype(dtype).itemsize:
            raise ValueError(f"{name}: {d.nbytes} bytes, need {count * np.dtype(dtype).itemsize}")
        return d.ptr

    def compute_forces_device(self, n: int, d_pos_fixed, d_species, d_acc, stream=None):
        n = int(n)
        self._check(self.lib.jtp_compute_forces(self.h, n, self._arg(d_pos_fixed, np.uint32, 3 * n, "pos"),
                                                self._arg(d_species, np.int8, n, "species"),
                                                self._arg(d_acc, self.real, 3 * n, "acc"), stream))

    def step_device(self, n: int, d_pos_fixed, d_vel, d_species, d_acc, dt: float, hubble_plus: float = 0.0,
                    hubble_minus: float | None = None, dtau_dt: float = 1.0, scheme: int = STEP_DKD,
                    graph: bool = False, stream=None):
        """One integrator step in place on device arrays (jtp_step): DKD is step_treepm_gpu's order,
        KDK needs d_acc = a(t) on entry.  hubble_minus defaults to hubble_plus (the reference's one H)."""
        n = int(n)
        hm = hubble_plus if hubble_minus is None else hubble_minus
        flags = int(scheme) | (STEP_GRAPH if graph else 0)
        self._check(self.lib.jtp_step(self.h, n, self._arg(d_pos_fixed, np.uint32, 3 * n, "pos"),
                                      self._arg(d_vel, self.real, 3 * n, "vel"),
                                      self._arg(d_species, np.int8, n, "species"),
                                      self._arg(d_acc, self.real, 3 * n, "acc"),
                                      flags, float(dt), float(hubble_plus), float(hm), float(dtau_dt), stream))

    def diagnostics(self, n: int, d_vel, d_species, stream=None) -> dict:
        """FP64 device reductions: kinetic energy and momentum per species, mesh potential energy."""
        ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
        d = Diag()
        self._check(self.lib.jtp_diagnostics(self.h, int(n), ptr(d_vel), ptr(d_species), C.byref(d), stream))
        return {"kinetic": np.array(d.kinetic[:]), "momentum": np.array([list(r) for r in d.momentum]),
                "pm_energy": d.pm_energy}

    def potential_energy(self, n: int, d_pos_fixed, d_species, stream=None) -> float:
        """FP64 total potential energy (FP64 context, erfc + Gaussian split): mesh + short range, mesh
        self-energy removed."""
        ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
        w = C.c_double(0.0)
        self._check(self.lib.jtp_potential_energy(self.h, int(n), ptr(d_pos_fixed), ptr(d_species), C.byref(w),
                                                  stream))
        return w.value

    def segregation(self, n: int, d_pos_fixed, d_species, n_cells: int, stream=None) -> tuple:
        """(Diff/Pois, corr(delta+, delta-), excess variance of m-) on an n_cells^3 count grid."""
        ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
        out = (C.c_double * 3)()
        self._check(self.lib.jtp_segregation(self.h, int(n), ptr(d_pos_fixed), ptr(d_species), int(n_cells), out,
                                             stream))
        return tuple(out)

    def generate_ics(self, params: "IcParams", d_pos_fixed, d_vel, d_species, vel_f64: bool = False,
                     stream=None) -> int:
        """Zel'dovich lattice ICs of n_grid^3 particles into device arrays; returns the m+ count."""
        ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
        n_plus = C.c_int64(0)
        self._check(self.lib.jtp_generate_ics(self.h, C.byref(params), ptr(d_pos_fixed), ptr(d_vel), int(vel_f64),
                                              ptr(d_species), C.byref(n_plus), stream))
        return n_plus.value

    def compute_forces(self, pos_fixed: np.ndarray, species: np.ndarray) -> np.ndarray:
        """Host convenience: upload, compute, download (float32 or float64 [n,3])."""
        n = len(species)
        dp = DeviceArray.from_host(np.ascontiguousarray(pos_fixed, np.uint32).reshape(n, 3))
        ds = DeviceArray.from_host(np.ascontiguousarray(species, np.int8))
        da = DeviceArray((n, 3), self.real)
        self.compute_forces_device(n, dp, ds, da)
        out = da.download()
        self._last = (dp, ds, da)
        return out

    def stats(self) -> dict:
        s = Stats()
        self._check(self.lib.jtp_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def stage(self, stage: int) -> np.ndarray:
        nbytes = C.c_size_t(0)
        self._check(self.lib.jtp_stage_size(self.h, stage, C.byref(nbytes)))
        buf = np.empty(nbytes.value, np.uint8)
        self._check(self.lib.jtp_stage_copy(self.h, stage, buf.ctypes.data, nbytes.value))
        rt = self.real
        view = {
            STAGE_KEYS: (np.uint64, (-1,)), STAGE_PERM: (np.int32, (-1,)), STAGE_NODES: (np.int32, (-1, 4)),
            STAGE_MONO: (np.float64, (-1, 8)), STAGE_RHO: (rt, (-1,)), STAGE_PHI: (rt, (-1,)),
            STAGE_ACC_PM: (rt, (-1, 3)), STAGE_SUMS: (np.uint64, (-1, 8)), STAGE_BOXES: (np.uint32, (-1, 6)),
            STAGE_RHO2: (rt, (-1,)), STAGE_PHI2: (rt, (-1,)),
        }[stage]
        return buf.view(view[0]).reshape(view[1])

    def pm_solve_device(self, d_rho, d_phi, stream=None):
        self._check(self.lib.jtp_pm_solve(self.h, d_rho.ptr, d_phi.ptr, stream))

    def to_fixed_device(self, n: int, d_pos, d_fixed, stream=None):
        if d_pos.dtype == np.float64:
            self._check(self.lib.jtp_pos_to_fixed_f64(self.h, int(n), d_pos.ptr, d_fixed.ptr, stream))
        else:
            self._check(self.lib.jtp_pos_to_fixed_f32(self.h, int(n), d_pos.ptr, d_fixed.ptr, stream))

    def close(self):
        if getattr(self, "h", None):
            self.lib.jtp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """A batch of independent boxes (jtp_ensemble_*): one parameter set and species masses per box."""

    def __init__(self, params: list, m_plus: list, m_minus_abs: list, n_streams: int = 4):
        self.lib = load_library()
        nb = len(params)
        arr = (Params * nb)(*params)
        mp = (C.c_double * nb)(*[float(x) for x in m_plus])
        mm = (C.c_double * nb)(*[float(x) for x in m_minus_abs])
        h = C.c_void_p()
        rc = self.lib.jtp_ensemble_create(nb, arr, mp, mm, int(n_streams), C.byref(h))
        if rc != JTP_OK:
            raise JtpError(rc, "jtp_ensemble_create failed")
        self.h, self.n_boxes = h, nb
        self.real = [np.float64 if p.precision == FP64 else np.float32 for p in params]

    def _check(self, rc):
        if rc != JTP_OK:
            raise JtpError(rc, self.lib.jtp_ensemble_last_error(self.h).decode())

    @staticmethod
    def _arg(d, dtype, count, name):
        """Raw pointer of a device argument; a DeviceArray must hold `count` elements of `dtype`."""
        if not isinstance(d, DeviceArray):
            return d
        if d.dtype != np.dtype(dtype):
            raise ValueError(f"{name}: dtype {d.dtype}, expected {np.dtype(dtype)}")
        if d.nbytes < count * np.dtype(dtype).itemsize:
            raise ValueError(f"{name}: {d.nbytes} bytes, need {count * np.dtype(dtype).itemsize}")
        return d.ptr

    def compute_forces_device(self, n, d_pos, d_species, d_acc, sequential=False, graph=False, per_box_walk=False,
                              stream=None):
        nb = self.n_boxes
        if not (len(n) == len(d_pos) == len(d_species) == len(d_acc) == nb):
            raise ValueError(f"expected {nb} entries per argument")
        n = [int(x) for x in n]
        nn = (C.c_int64 * nb)(*n)
        pp = (C.c_void_p * nb)(*[self._arg(d, np.uint32, 3 * max(k, 0), f"pos[{b}]") for b, (d, k) in enumerate(zip(d_pos, n))])
        ss = (C.c_void_p * nb)(*[self._arg(d, np.int8, max(k, 0), f"species[{b}]")
                                 for b, (d, k) in enumerate(zip(d_species, n))])
        aa = (C.c_void_p * nb)(*[self._arg(d, self.real[b], 3 * max(k, 0), f"acc[{b}]")
                                 for b, (d, k) in enumerate(zip(d_acc, n))])
        flags = ((ENS_SEQUENTIAL if sequential else 0) | (ENS_GRAPH if graph else 0)
                 | (ENS_WALK_PER_BOX if per_box_walk else 0))
        self._check(self.lib.jtp_ensemble_compute_forces(self.h, nn, pp, ss, aa, flags, stream))

    def stats(self, box: int = -1) -> dict:
        st = Stats()
        self._check(self.lib.jtp_ensemble_get_stats(self.h, int(box), C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.jtp_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_fixed(x: np.ndarray, box_size: float) -> np.ndarray:
    """Host conversion of positions in [-L/2, L/2) (any period) to 32-bit fixed point."""
    t = np.asarray(x, np.float64) / box_size + 0.5
    t = t - np.floor(t)
    w = (t * 4294967296.0).astype(np.uint64)
    w[w >= 4294967296] -= 4294967296
    return w.astype(np.uint32)


def from_fixed(u: np.ndarray, box_size: float) -> np.ndarray:
    return (np.asarray(u, np.float64) / 4294967296.0 - 0.5) * box_size


# ---------------------------------------------------------------- reference-shaped facades
@dataclass
class Particle:
    """nbody.rs Particle (pos, vel, |mass|, sign); only pos / mass / sign matter for forces."""
    pos: tuple
    mass: float
    sign: int


class TreePMForce:
    """Mirror of TreePMForce (treepm_force.rs:17-146) on the GPU.

    ``TreePMForce(r_cut, grid_size, box_size, theta, softening)`` keeps the
    reference constructor; the split scale defaults to r_cut/3 like the
    reference (treepm_force.rs:36) unless ``r_s`` is given.  Positions are in
    [-L/2, L/2) like the reference; species masses are per species (the
    reference's per-particle |m| must be uniform within a species).
    """

    def __init__(self, r_cut, grid_size, box_size, theta, softening, r_s=None, precision=FP64,
                 split_mode=SPLIT_ERFC_GAUSS, gradient=GRAD_4PT, deconvolve=1, g_constant=1.0, pm_only=False,
                 interlace=1):
        p = default_params(box_size, grid_size)
        p.r_cut = float(r_cut)
        p.r_s = float(r_cut / 3.0 if r_s is None else r_s)
        p.theta = float(theta)
        p.softening = float(softening)
        p.precision = precision
        p.split_mode = split_mode
        p.gradient = gradient
        p.deconvolve = deconvolve
        p.g_constant = g_constant
        p.interlace = interlace
        p.tree_enabled = 0 if pm_only else 1
        self.params = p
        self.ctx = ForceContext(p)
        self._acc = None

    @classmethod
    def new_pm_only(cls, grid_size, box_size):
        return cls(box_size / 8.0, grid_size, box_size, 0.5, 0.5, r_s=0.0, pm_only=True, split_mode=SPLIT_POLY4)

    def update(self, particles):
        pos = np.array([p.pos for p in particles], np.float64).reshape(-1, 3)
        sign = np.array([1 if p.sign > 0 else -1 for p in particles], np.int8)
        mass = np.array([abs(p.mass) for p in particles], np.float64)
        mp = mass[sign > 0]
        mm = mass[sign < 0]
        m_plus = float(mp[0]) if len(mp) else 1.0
        m_minus = float(mm[0]) if len(mm) else 1.0
        if (len(mp) and not np.all(mp == m_plus)) or (len(mm) and not np.all(mm == m_minus)):
            raise JtpError(JTP_ERR_ARG, "masses must be uniform within a species")
        self.ctx.set_species_mass(m_plus, m_minus)
        self._acc = self.ctx.compute_forces(to_fixed(pos, self.params.box_size), sign).astype(np.float64)

    def compute_all_forces(self, particles=None):
        if self._acc is None:
            raise JtpError(JTP_ERR_STATE, "update() must be called first")
        return self._acc.copy()


class PoissonSolver:
    """Mirror of CuFFTPoisson / solve_device (cufft_ffi.rs:71-337) for device grids."""

    def __init__(self, grid_size, box_size, g_constant=1.0, r_s=0.0, k_min=0, precision=FP64, deconvolve=0):
        p = default_params(box_size, grid_size)
        p.r_s = float(r_s)
        p.k_min = int(k_min)
        p.g_constant = float(g_constant)
        p.precision = precision
        p.deconvolve = deconvolve
        p.split_mode = SPLIT_POLY4 if r_s <= 0 else SPLIT_ERFC_GAUSS
        p.r_cut = box_size / 8.0
        self.ctx = ForceContext(p)
        self.grid_size = grid_size

    def solve_noalloc_device(self, d_rho, d_phi, stream=None):
        """The same solve through jtp_pm_solve_noalloc (no context buffers, no allocation; asynchronous)."""
        ptr = lambda d: d.ptr if isinstance(d, DeviceArray) else d
        rc = load_library().jtp_pm_solve_noalloc(C.byref(self.ctx.params), ptr(d_rho), ptr(d_phi), stream)
        if rc != JTP_OK:
            raise JtpError(rc, "jtp_pm_solve_noalloc failed")

    def solve(self, rho: np.ndarray) -> np.ndarray:
        """rho: mass per cell, [G,G,G] indexed [z][y][x]; returns the potential."""
        d_rho = DeviceArray.from_host(np.ascontiguousarray(rho, self.ctx.real))
        d_phi = DeviceArray(rho.shape, self.ctx.real)
        self.ctx.pm_solve_device(d_rho, d_phi)
        return d_phi.download()
