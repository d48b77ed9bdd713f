"""Segregation metrics (compute_segregation_metrics, src/bin/janus_mu19_production.rs:219-273).

CPU: the oracle against closed forms that fix each of the three outputs, and the fixed-point binning
against the reference's f32 binning.  GPU: jtp_segregation against the oracle (counts are exact, the
moments agree to FP64 rounding) and against the same closed forms."""
import numpy as np
import pytest


def _cell_centres(nc, L, cells):
    """Positions in [-L/2, L/2) at the centres of the given (ix, iy, iz) cells."""
    h = L / nc
    return (np.asarray(cells, np.float64) + 0.5) * h - L / 2


def _segregated(nc, L, k):
    """k m+ in every cell with even ix + iy + iz, k m- in every odd one."""
    I, J, K = np.meshgrid(*(np.arange(nc),) * 3, indexing="ij")
    cells = np.stack([I, J, K], -1).reshape(-1, 3)
    par = cells.sum(1) % 2
    pos = np.repeat(_cell_centres(nc, L, cells), k, axis=0)
    sp = np.repeat(np.where(par == 0, 1, -1), k).astype(np.int8)
    return pos, sp


def _identical(nc, L, seed):
    """m+ and m- at the same random positions (every cell holds as many of each)."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, nc, (3000, 3))
    pos = _cell_centres(nc, L, cells)
    return np.concatenate([pos, pos]), np.concatenate([np.ones(3000), -np.ones(3000)]).astype(np.int8)


def test_oracle_closed_forms(orc):
    L, nc = 100.0, 8
    pos, sp = _segregated(nc, L, 3)
    d, c, e = orc.segregation(orc.fixed(pos, L), sp, nc)
    # n+ - n- = +-k around 0: Var = k^2, Poisson k; delta+ = -delta- = +-1; Var(n-) = k^2/4, mean k/2
    assert d == pytest.approx(3.0, rel=1e-12)
    assert c == pytest.approx(-1.0, rel=1e-12)
    assert e == pytest.approx(3 / 2 - 1, rel=1e-12)
    pos, sp = _identical(nc, L, 1)
    d, c, e = orc.segregation(orc.fixed(pos, L), sp, nc)
    assert d == 0.0 and c == pytest.approx(1.0, rel=1e-12)
    counts = np.bincount(np.ravel_multi_index(((pos[:3000] + L / 2) // (L / nc)).astype(int).T, (nc,) * 3),
                         minlength=nc ** 3)
    assert e == pytest.approx((counts.var() - counts.mean()) / counts.mean(), rel=1e-12)
    # one of each per cell: no variance, correlation undefined (reported 0), excess variance -1
    I, J, K = np.meshgrid(*(np.arange(nc),) * 3, indexing="ij")
    p1 = _cell_centres(nc, L, np.stack([I, J, K], -1).reshape(-1, 3))
    d, c, e = orc.segregation(orc.fixed(np.concatenate([p1, p1]), L),
                              np.repeat([1, -1], len(p1)).astype(np.int8), nc)
    assert (d, c, e) == (0.0, 0.0, -1.0)


def test_fixed_binning_matches_reference_binning(orc):
    """Away from cell faces the fixed-point bins are the reference's f32 bins (non-power-of-two grid)."""
    rng = np.random.default_rng(3)
    L, nc = 250.0, 37
    h = L / nc
    cells = rng.integers(0, nc, (20000, 3))
    pos = (cells + rng.uniform(0.05, 0.95, (20000, 3))) * h - L / 2
    sp = rng.choice([1, -1], 20000).astype(np.int8)
    a = orc.segregation(orc.fixed(pos, L), sp, nc)
    b = orc.segregation_f32_ref(pos.astype(np.float32), sp, L, nc)
    assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [64, 50])
def test_gpu_matches_oracle(jtp, orc, nc):
    print(f"case: segregation vs oracle n_cells={nc}", flush=True)
    from janus_treepm import synth
    w = synth.clustered(n_total=300_000, n_cells=64, n_halos=60, seed=21)
    pos = jtp.to_fixed(w.pos, w.box_size)
    ctx = jtp.ForceContext(jtp.default_params(w.box_size, 64))
    dp, ds = jtp.DeviceArray.from_host(pos), jtp.DeviceArray.from_host(w.species)
    got = np.array(ctx.segregation(w.n, dp, ds, nc))
    ref = orc.segregation(pos, w.species, nc)
    print("gpu", got, "oracle", ref)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-13)
    assert np.array_equal(np.array(ctx.segregation(w.n, dp, ds, nc)), got)  # bit-reproducible


@pytest.mark.gpu
def test_gpu_closed_forms(jtp):
    print("case: segregation closed forms", flush=True)
    L, nc = 100.0, 8
    ctx = jtp.ForceContext(jtp.default_params(L, 64))
    pos, sp = _segregated(nc, L, 3)
    d, c, e = ctx.segregation(len(sp), jtp.DeviceArray.from_host(jtp.to_fixed(pos, L)), jtp.DeviceArray.from_host(sp), nc)
    assert d == pytest.approx(3.0, rel=1e-12) and c == pytest.approx(-1.0, rel=1e-12)
    assert e == pytest.approx(0.5, rel=1e-12)
    with pytest.raises(jtp.JtpError):
        ctx.segregation(len(sp), jtp.DeviceArray.from_host(jtp.to_fixed(pos, L)), jtp.DeviceArray.from_host(sp), 0)
