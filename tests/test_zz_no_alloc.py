"""jtp_pm_solve_noalloc (the allocation-free Poisson solve) against jtp_pm_solve on a context with the same
parameters: two calls in a row on the same GPU, the second on a smaller mesh (the scratch array keeps the
first call's contents), both precisions, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [1, 0])
def test_noalloc_solve_matches_context_solve(jtp, precision):
    rt = np.float64 if precision == 1 else np.float32
    tol = 1e-12 if precision == 1 else 2e-6
    for G, k_min, seed in ((128, 0, 3), (32, 2, 4)):  # second call smaller than the first
        print(f"case: noalloc solve precision={precision} G={G} k_min={k_min}", flush=True)
        L = 100.0
        ps = jtp.PoissonSolver(G, L, g_constant=1.0, r_s=96.0 / G, k_min=k_min, precision=precision, deconvolve=1)
        rho = np.random.default_rng(seed).standard_normal((G, G, G)).astype(rt)
        ref = ps.solve(rho).astype(np.float64)
        d_rho = jtp.DeviceArray.from_host(rho)
        d_phi = jtp.DeviceArray((G, G, G), rt)
        d_phi.upload(np.full((G, G, G), np.nan, rt))
        ps.solve_noalloc_device(d_rho, d_phi)
        assert jtp.load_library().jtp_device_synchronize() == jtp.JTP_OK
        phi = d_phi.download().astype(np.float64)
        err = np.max(np.abs(phi - ref)) / np.max(np.abs(ref))
        print(f"  max |noalloc - context| / max |phi| = {err:.2e}", flush=True)
        assert np.all(np.isfinite(phi)) and err <= tol


def test_noalloc_solve_rejects_large_mesh(jtp):
    print("case: noalloc size limit", flush=True)
    lib = jtp.load_library()
    p = jtp.default_params(100.0, 256)
    d = jtp.DeviceArray((8,), np.float32)
    assert lib.jtp_pm_solve_noalloc(C.byref(p), d.ptr, d.ptr, None) == jtp.JTP_ERR_ARG
    p = jtp.default_params(100.0, 64)
    assert lib.jtp_pm_solve_noalloc(C.byref(p), None, d.ptr, None) == jtp.JTP_ERR_ARG
    assert lib.jtp_pm_solve_noalloc(None, d.ptr, d.ptr, None) == jtp.JTP_ERR_ARG
