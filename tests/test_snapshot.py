"""Snapshot I/O in the reference's formats (include/janus_treepm.h, csrc/jtp_io.hip).

The expected bytes are laid out here with numpy from the reference writers' field order:
production records, src/bin/janus_mu19_production.rs:199-216 (u32 N, f32 L, u32 step, f32 z; per
particle 6 f32 + i8); v3, src/snapshot_v3.rs:108-157 (header) and :313-335 (36-byte records).
CPU: header parsing of files built that way, and the reference's header tests (test_roundtrip_header,
test_header_size_field, bad magic / version).  GPU: written files equal the expected bytes exactly;
reading them back restores the device arrays (positions to the f32 resolution of the file, through the
oracle's fixed-point conversion); several staging chunks and ragged tails; invalid sign bytes."""
import os

import numpy as np
import pytest

MAGIC = 0x0A335653554E414A
V3_HEAD = np.dtype([
    ("magic", "<u8"), ("version", "<u4"), ("header_size", "<u4"), ("n_total", "<u8"), ("a", "<f8"),
    ("t_gyr", "<f8"), ("l_box", "<f8"), ("h0", "<f8"), ("mu", "<f8"), ("omega_b", "<f8"),
    ("m_part_plus_base", "<f8"), ("m_part_minus_base", "<f8"), ("eps_plus_base", "<f8"),
    ("eps_minus_base", "<f8"), ("n_split_max", "<u4"), ("seed_ic", "<u4"), ("z_init", "<f8"),
    ("n_stars", "<u8"), ("z_start_run", "<f8"), ("sfr", "<f8"), ("rho_max", "<f8"), ("run_label", "S256"),
])
V3_REC = np.dtype([("pos", "<f4", 3), ("vel", "<f4", 3), ("mass", "<f4"), ("epsilon", "<f4"), ("sign", "u1"),
                   ("split_level", "u1"), ("is_star", "u1"), ("flags", "u1")])
PROD_HEAD = np.dtype([("n", "<u4"), ("box", "<f4"), ("step", "<u4"), ("z", "<f4")])
PROD_REC = np.dtype([("pos", "<f4", 3), ("vel", "<f4", 3), ("sign", "i1")])


def test_layouts():
    assert V3_HEAD.itemsize == 408 and V3_REC.itemsize == 36
    assert PROD_HEAD.itemsize == 16 and PROD_REC.itemsize == 25


def _lib():
    import janus_treepm as j
    return j


def _v3_header(n, **kw):
    h = np.zeros(1, V3_HEAD)
    h["magic"], h["version"], h["header_size"], h["n_total"] = MAGIC, 3, 408, n
    fields = dict(a=0.5, t_gyr=3.25, l_box=500.0, h0=69.9, mu=19.0, omega_b=0.05, m_part_plus_base=1e10,
                  m_part_minus_base=2e10, eps_plus_base=0.5, eps_minus_base=2.5, n_split_max=2, seed_ic=42,
                  z_init=10.0, n_stars=7, z_start_run=9.0, sfr=1.5, rho_max=3e14, run_label=b"test_label")
    fields.update(kw)
    for k, v in fields.items():
        h[k] = v
    return h


def test_v3_header_roundtrip_fields(tmp_path):
    """test_roundtrip_header (snapshot_v3.rs:545-627): every header field survives."""
    j = _lib()
    path = str(tmp_path / "h.v3")
    h = _v3_header(0)
    open(path, "wb").write(h.tobytes())
    m = j.read_snapshot_meta(path, j.SNAP_V3)
    assert m.n == 0 and m.box_size == 500.0 and m.a == 0.5 and m.t_gyr == 3.25 and m.h0 == 69.9
    assert m.mu == 19.0 and m.omega_b == 0.05 and m.m_part_plus_base == 1e10 and m.m_part_minus_base == 2e10
    assert m.eps_plus_base == 0.5 and m.eps_minus_base == 2.5 and m.n_split_max == 2 and m.seed_ic == 42
    assert m.z_init == 10.0 and m.n_stars == 7 and m.z_start_run == 9.0 and m.sfr == 1.5 and m.rho_max == 3e14
    assert m.run_label == b"test_label" and m.z == pytest.approx(1.0)


def test_v3_header_rejects_bad_magic_and_version(tmp_path):
    j = _lib()
    for field, val in (("magic", 0x1234), ("version", 2), ("header_size", 400)):
        h = _v3_header(0)
        h[field] = val
        path = str(tmp_path / f"bad_{field}.v3")
        open(path, "wb").write(h.tobytes())
        with pytest.raises(j.JtpError):
            j.read_snapshot_meta(path, j.SNAP_V3)


def test_production_header(tmp_path):
    j = _lib()
    h = np.zeros(1, PROD_HEAD)
    h["n"], h["box"], h["step"], h["z"] = 5, 250.0, 1234, 0.5
    path = str(tmp_path / "p.snap")
    open(path, "wb").write(h.tobytes())
    m = j.read_snapshot_meta(path, j.SNAP_PRODUCTION)
    assert (m.n, m.box_size, m.step, m.z) == (5, 250.0, 1234, 0.5)
    with pytest.raises(j.JtpError):
        j.read_snapshot_meta(str(tmp_path / "missing"), j.SNAP_PRODUCTION)


def test_meta_defaults():
    """SnapshotHeaderV3::default (snapshot_v3.rs:48-76)."""
    j = _lib()
    m = j.snap_meta(j.SNAP_V3, 500.0)
    assert (m.a, m.h0, m.mu, m.omega_b, m.eps_plus_base, m.eps_minus_base, m.seed_ic, m.z_init) == \
        (1.0, 69.9, 19.0, 0.05, 0.5, 2.5, 42, 10.0)
    assert m.velocity_scale == 1.0


def _particles(n, L, seed):
    from janus_treepm import to_fixed
    rng = np.random.default_rng(seed)
    pos = to_fixed(rng.uniform(-L / 2, L / 2, (n, 3)), L)
    vel = rng.normal(0, 300.0, (n, 3)).astype(np.float32)
    sp = rng.choice(np.array([1, -1], np.int8), n)
    return pos, vel, sp


def _f32_pos(pos, L):
    return ((pos.astype(np.float64) / 4294967296.0 - 0.5) * L).astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,n", [(1, 1000), (3, 1000), (1, (1 << 20) + 777), (3, (1 << 20) + 3), (1, 0)])
def test_write_bytes_and_read_back(jtp, orc, tmp_path, fmt, n):
    print(f"case: snapshot format={fmt} n={n}", flush=True)
    L, vs = 250.0, 2.5
    pos, vel, sp = _particles(n, L, 5)
    dp, dv, ds = (jtp.DeviceArray.from_host(a) for a in (pos, vel, sp))
    meta = jtp.snap_meta(fmt, L, n=n, step=77, z=1.5, velocity_scale=vs, run_label="roundtrip",
                         mass_plus=1.0, mass_minus=19.0, eps_plus=0.1, eps_minus=0.3)
    path = str(tmp_path / "s.bin")
    jtp.write_snapshot(path, meta, dp, dv, ds)
    raw = open(path, "rb").read()
    fpos = _f32_pos(pos, L)
    fvel = vel * np.float32(vs)
    if fmt == jtp.SNAP_PRODUCTION:
        h = np.zeros(1, PROD_HEAD)
        h["n"], h["box"], h["step"], h["z"] = n, L, 77, 1.5
        r = np.zeros(n, PROD_REC)
        r["pos"], r["vel"], r["sign"] = fpos, fvel, sp
    else:
        h = _v3_header(n, a=1.0, t_gyr=0.0, l_box=L, m_part_minus_base=1e10, n_split_max=0, n_stars=0,
                       z_start_run=10.0, sfr=0.0, rho_max=0.0, run_label=b"roundtrip")
        r = np.zeros(n, V3_REC)
        r["pos"], r["vel"] = fpos, fvel
        r["mass"] = np.where(sp > 0, 1.0, 19.0)
        r["epsilon"] = np.where(sp > 0, 0.1, 0.3).astype(np.float32)
        r["sign"] = np.where(sp > 0, 1, 255)
    assert raw == h.tobytes() + r.tobytes()
    # read back into fresh arrays
    m = jtp.read_snapshot_meta(path, fmt)
    assert m.n == n
    rp = jtp.DeviceArray((n, 3), np.uint32)
    rv = jtp.DeviceArray((n, 3), np.float64)
    rs = jtp.DeviceArray((n,), np.int8)
    jtp.read_snapshot(path, fmt, n, rp, rv, rs, vel_f64=True, velocity_scale=vs)
    if n:
        assert np.array_equal(rp.download(), orc.fixed(fpos.astype(np.float64), L))
        assert np.array_equal(rv.download(), (fvel * np.float32(1 / vs)).astype(np.float64))
        assert np.array_equal(rs.download(), sp)
        # the file's f32 positions are within an f32 half-ulp of the originals
        err = np.abs(jtp.from_fixed(rp.download(), L) - jtp.from_fixed(pos, L))
        assert err.max() <= L / 2 * 2.0 ** -24


@pytest.mark.gpu
def test_v3_header_size_skip_and_bad_sign(jtp, tmp_path):
    """test_header_size_field (snapshot_v3.rs:664-707): a larger header_size is skipped over."""
    print("case: snapshot v3 header_size and sign checks", flush=True)
    n, L = 300, 100.0
    h = _v3_header(n, l_box=L, header_size=416)
    r = np.zeros(n, V3_REC)
    rng = np.random.default_rng(2)
    r["pos"] = rng.uniform(-L / 2, L / 2, (n, 3)).astype(np.float32)
    r["vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    r["sign"] = np.where(np.arange(n) % 3 == 0, 255, 1)
    path = str(tmp_path / "big_header.v3")
    open(path, "wb").write(h.tobytes() + b"\xab" * 8 + r.tobytes())
    dp, dv, ds = jtp.DeviceArray((n, 3), np.uint32), jtp.DeviceArray((n, 3), np.float32), jtp.DeviceArray((n,), np.int8)
    jtp.read_snapshot(path, jtp.SNAP_V3, n, dp, dv, ds)
    assert np.array_equal(dv.download(), r["vel"])
    assert np.array_equal(ds.download(), np.where(r["sign"] == 1, 1, -1).astype(np.int8))
    with pytest.raises(jtp.JtpError):  # wrong count
        jtp.read_snapshot(path, jtp.SNAP_V3, n - 1, dp, dv, ds)
    r["sign"][17] = 7
    open(path, "wb").write(h.tobytes() + b"\xab" * 8 + r.tobytes())
    with pytest.raises(jtp.JtpError):
        jtp.read_snapshot(path, jtp.SNAP_V3, n, dp, dv, ds)
    open(path, "wb").write(h.tobytes() + b"\xab" * 8 + r.tobytes()[:-36])
    with pytest.raises(jtp.JtpError):  # truncated
        jtp.read_snapshot(path, jtp.SNAP_V3, n, dp, dv, ds)
    os.remove(path)
