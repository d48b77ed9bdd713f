"""TreePM particle-force-updates/s on one MI355X: FP32 (fixed-point) form vs the plain FP64 form.

Workload: the clustered two-species box (N = 2.2e7, 1.1e7 m+ / 1.1e7 m-, 60 % of m+ in 2000 NFW
halos, m- rejection-sampled in under-dense regions), L = 250, PM 256^3, r_s = 1.25 cells,
r_cut = 4.5 r_s, theta = 0.5, leaf 4, m+ = 1, |m-| = 19, eps = L / N^(1/3) / 30.
A step is one complete force evaluation (sort, tree, deposit, FFT, interpolation, walk) on inputs
already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "janus-sim_amd"))
sys.path.insert(0, ROOT)

import janus_treepm as j  # noqa: E402
from janus_treepm import autotune, synth  # noqa: E402

METRIC = "TreePM particle-force-updates/s at N=2.2e7 on 1 GPU, FP32 vs FP64 at matched error"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOPS_PER_INTERACTION = 38


def log(msg):
    print(msg, file=sys.stderr, flush=True)


OVERRIDES = {}


def make_ctx(w, precision):
    p = j.default_params(w.box_size, w.n_cells)
    p.r_s, p.r_cut, p.theta, p.softening, p.leaf_size = w.r_s, w.r_cut, w.theta, w.softening, w.leaf_size
    p.precision = precision
    for k, v in OVERRIDES.items():
        setattr(p, k, v)
    return j.ForceContext(p, w.m_plus, w.m_minus_abs)


def run_form(w, d_pos, d_sp, precision, warmup, steps):
    ctx = make_ctx(w, precision)
    rt = np.float64 if precision == j.FP64 else np.float32
    d_acc = j.DeviceArray((w.n, 3), rt)
    for i in range(warmup):
        ctx.compute_forces_device(w.n, d_pos, d_sp, d_acc)
        log(f"  warmup {i + 1}/{warmup} precision={'fp64' if precision else 'fp32'}")
    j.load_library().jtp_device_synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.compute_forces_device(w.n, d_pos, d_sp, d_acc)
        stats.append(ctx.stats())
    j.load_library().jtp_device_synchronize()
    t = time.perf_counter() - t0
    log(f"  timed {steps} steps precision={'fp64' if precision else 'fp32'}: {t / steps * 1e3:.2f} ms/step")
    ctx.close()
    return t, stats, d_acc


def phase_summary(stats):
    keys = ["ms_sort", "ms_build", "ms_deposit", "ms_fft", "ms_interp", "ms_walk", "ms_total"]
    return {k: float(np.mean([s[k] for s in stats])) for k in keys}


def cpu_baseline(w_full):
    """Reference CPU TreePM (oracle restatement of treepm_force.rs / tree_short.rs / pm_grid.rs: leaf-1 octree,
    x^4 split, r_s = r_cut/3, two-grid PM) timed on a bounded sample (about 10 s on 16 cores): the same
    recipe in a box of 1/8 the volume at the same number density, cell size, r_s, theta and softening."""
    from oracle.oracle_lib import RefCpu, load

    orc = load()
    n_s = w_full.n // 8
    Ls = w_full.box_size / 2
    Gs = w_full.n_cells // 2
    ws = synth.clustered(n_total=n_s, box_size=Ls, n_cells=Gs, n_halos=max(1, 2000 // 8), seed=99)
    h = Ls / Gs
    r_s = w_full.r_s
    r_cut = 3.0 * r_s
    mass = np.where(ws.species > 0, ws.m_plus, ws.m_minus_abs)
    ref = RefCpu(orc, r_cut, Gs, Ls, w_full.theta, w_full.softening, r_s, 1.0 / h ** 3, 1.0)
    t0 = time.perf_counter()
    ref.update(ws.pos, mass, ws.species)
    ref.forces()
    t = time.perf_counter() - t0
    return {
        "value": n_s / t, "unit": "particle-force-updates/s", "cores": int(orc.lib.orc_num_threads()),
        "kind": "port",
        "sample": f"C++ restatement of the reference CPU TreePM (leaf-1 octree, x^4 split, r_s=r_cut/3, two-grid "
                  f"FP64 PM) on the same recipe at 1/8 volume: N={n_s}, L={Ls:g}, PM {Gs}^3, theta={w_full.theta}, "
                  f"eps={w_full.softening:.4g}; one update + all forces in {t:.2f} s",
    }


def config0(warmup=2, steps=10):
    """The small configuration timed in full: N = 2 x 64^3 lattice + Zel'dovich (P(k) ~ k^-2.5, rms 0.3
    cell), L = 250, PM 64^3, r_s = 1.25 cells, r_cut = 4.5 r_s, theta = 0.5, leaf 8, m+ = 1, |m-| = 19:
    one force evaluation on the GPU in both forms (inputs resident in HBM), and one update + all forces
    of the C++ restatement of the reference CPU TreePM (r_s = r_cut / 3 with r_s = 1.25 cells) on the
    whole particle set, all host threads."""
    from oracle.oracle_lib import RefCpu, load

    w = synth.lattice_zeldovich(n_side=64, box_size=250.0, n_cells=64, seed=42, leaf_size=8)
    pos = j.to_fixed(w.pos, w.box_size)
    d_pos, d_sp = j.DeviceArray.from_host(pos), j.DeviceArray.from_host(w.species)
    out = {"n_particles": w.n, "pm_grid": w.n_cells, "box": w.box_size, "leaf_size": 8}
    for prec, label in ((j.FP64, "fp64"), (j.FP32, "fp32")):
        t, _, d_acc = run_form(w, d_pos, d_sp, prec, warmup, steps)
        d_acc.free()
        out[label] = {"ms_per_eval": t / steps * 1e3, "updates_per_s": w.n * steps / t}
    orc = load()
    h = w.box_size / w.n_cells
    mass = np.where(w.species > 0, w.m_plus, w.m_minus_abs)
    ref = RefCpu(orc, 3.0 * w.r_s, w.n_cells, w.box_size, w.theta, w.softening, w.r_s, 1.0 / h ** 3, 1.0)
    t0 = time.perf_counter()
    ref.update(w.pos, mass, w.species)
    ref.forces()
    t = time.perf_counter() - t0
    out["cpu_reference_port"] = {"ms_per_eval": t * 1e3, "updates_per_s": w.n / t,
                                  "cores": int(orc.lib.orc_num_threads())}
    out["fp64_gpu_over_cpu"] = out["cpu_reference_port"]["ms_per_eval"] / out["fp64"]["ms_per_eval"]
    return out


def measured_traffic(workload):
    """HBM bytes per walk launch from the committed FETCH_SIZE / WRITE_SIZE summary of this workload
    (profiles/walk_fp32_<workload>_hbm_traffic.txt, calibrated against a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", f"walk_fp32_{workload}_hbm_traffic.txt")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("MACHINE: "):
            return float(json.loads(line[len("MACHINE: "):])["traffic_bytes"]), os.path.relpath(path, ROOT)
    return None, None


WORKLOADS = {
    "clustered_22M": dict(n_total=22_000_000, box_size=250.0, n_cells=256, n_halos=2000, seed=7),
    "clustered_344k": dict(n_total=343_750, box_size=62.5, n_cells=64, n_halos=31, seed=99),
    # the multistep box's particles (2 x 256^3 + 7.2e6 m+, L = 300, PM 512^3) as a single force evaluation
    "clustered_41M": dict(n_total=256 ** 3 * 2 + 7_200_000, n_plus=256 ** 3 + 7_200_000, box_size=300.0,
                          n_cells=512, n_halos=8000, seed=11),
}


def accuracy(w, name, d_pos, d_sp, pos_fixed):
    """Median / p99 relative acceleration error per species of both forms against the committed FP64 Ewald
    truth for this workload (tests/golden/ewald_<name>.npz, made by tests/golden/make_ewald_truth.py)."""
    path = os.path.join(ROOT, "tests", "golden", f"ewald_{name}.npz")
    if not os.path.exists(path):
        return {"status": f"no truth fixture for {name}"}
    f = np.load(path)
    if str(f["digest"]) != w.digest(pos_fixed):
        return {"status": "inputs differ from the fixture's (digest mismatch)"}
    tg, truth = f["targets"], f["acc"]
    out = {"targets": int(len(tg)), "status": "ok"}
    for prec, label in ((j.FP32, "fp32"), (j.FP64, "fp64")):
        ctx = make_ctx(w, prec)
        d_acc = j.DeviceArray((w.n, 3), np.float64 if prec == j.FP64 else np.float32)
        ctx.compute_forces_device(w.n, d_pos, d_sp, d_acc)
        acc = d_acc.download()[tg].astype(np.float64)
        ctx.close()
        err = np.linalg.norm(acc - truth, axis=1) / np.linalg.norm(truth, axis=1)
        res = {}
        for sp, nm in ((1, "m+"), (-1, "m-")):
            e = err[w.species[tg] == sp]
            res[nm] = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99))}
        res["all"] = {"median": float(np.median(err)), "p99": float(np.percentile(err, 99))}
        out[label] = res
    return out


def run_ensemble(args):
    """Uneven ensemble of 48 boxes (synth.ensemble_c3): total particle-force-updates/s of the FP32 form
    with the boxes prepared sequentially on one stream, concurrently over --ensemble-streams streams
    (default 4) or as one replayed graph, and walked either by one launch over all boxes' groups (the
    default) or box by box ("per_box" modes)."""
    t0 = time.perf_counter()
    ws = synth.ensemble_c3()
    log(f"ensemble generated in {time.perf_counter() - t0:.1f} s ({len(ws)} boxes, N={sum(w.n for w in ws)})")
    params = []
    for w in ws:
        p = j.default_params(w.box_size, w.n_cells)
        p.r_s, p.r_cut, p.theta, p.softening, p.leaf_size = w.r_s, w.r_cut, w.theta, w.softening, w.leaf_size
        for k, v in OVERRIDES.items():
            setattr(p, k, v)
        params.append(p)
    ens = j.Ensemble(params, [w.m_plus for w in ws], [w.m_minus_abs for w in ws], n_streams=args.ensemble_streams)
    dpos = [j.DeviceArray.from_host(j.to_fixed(w.pos, w.box_size)) for w in ws]
    dsp = [j.DeviceArray.from_host(w.species) for w in ws]
    dacc = [j.DeviceArray((w.n, 3), np.float32) for w in ws]
    n = [w.n for w in ws]
    ntot = sum(n)
    modes = {}
    # baseline: the boxes one after another as separate force calls (each call synchronises and reads
    # its status back, as a host loop over boxes would)
    ctxs = [j.ForceContext(p, w.m_plus, w.m_minus_abs) for p, w in zip(params, ws)]
    for _ in range(args.warmup):
        for b, ctx in enumerate(ctxs):
            ctx.compute_forces_device(n[b], dpos[b], dsp[b], dacc[b])
    t = time.perf_counter()
    for _ in range(args.steps):
        for b, ctx in enumerate(ctxs):
            ctx.compute_forces_device(n[b], dpos[b], dsp[b], dacc[b])
    dt = (time.perf_counter() - t) / args.steps
    modes["separate_calls"] = {"ms_per_eval": dt * 1e3, "updates_per_s": ntot / dt}
    log(f"  separate calls: {dt * 1e3:.2f} ms per evaluation of all boxes")
    for ctx in ctxs:
        ctx.close()
    for mode in ("per_box_sequential", "per_box_concurrent", "per_box_graph", "sequential", "concurrent", "graph"):
        kw = dict(sequential=mode.endswith("sequential"), graph=mode.endswith("graph"),
                  per_box_walk=mode.startswith("per_box"))
        for _ in range(args.warmup):
            ens.compute_forces_device(n, dpos, dsp, dacc, **kw)
        j.load_library().jtp_device_synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            ens.compute_forces_device(n, dpos, dsp, dacc, **kw)
        j.load_library().jtp_device_synchronize()
        dt = (time.perf_counter() - t) / args.steps
        modes[mode] = {"ms_per_eval": dt * 1e3, "updates_per_s": ntot / dt}
        log(f"  ensemble {mode}: {dt * 1e3:.2f} ms per evaluation of all boxes")
    st = ens.stats()
    best = max((m for m in modes if m != "separate_calls"), key=lambda m: modes[m]["updates_per_s"])
    out = {
        "metric": "TreePM particle-force-updates/s over an uneven ensemble of 48 boxes on 1 GPU",
        "value": modes[best]["updates_per_s"], "unit": "particle-force-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": modes[best]["ms_per_eval"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"48 lattice+Zel'dovich boxes, N_b = 2 n^3 log2-uniform in [2*32^3, 2*128^3] "
                               f"(total {ntot}), G in {{32,64,128}}, L ~ U[100,500], |m-|/m+ ~ U[5,40]",
                   "best_mode": best, "parallelism": "single GPU", "overrides": OVERRIDES},
        "modes": modes, "interactions": st["interactions"],
        "best_over_separate_calls": modes["separate_calls"]["ms_per_eval"] / modes[best]["ms_per_eval"],
        "concurrent_over_sequential": modes["sequential"]["ms_per_eval"] / modes["concurrent"]["ms_per_eval"],
        "graph_over_sequential": modes["sequential"]["ms_per_eval"] / modes["graph"]["ms_per_eval"],
    }
    ens.close()
    print(json.dumps(out), flush=True)


def run_multistep(args):
    """Large single box, multi-step: N = 2 x 256^3 + 7.2e6 m+ (40,754,432), L = 300, PM 512^3, theta = 0.5,
    the clustered recipe with 8000 halos; 100 KDK steps of the FP32 form (hipGraph replay, FP64 drift)
    and the same run in the FP64 form, from rest.  Reports the per-step time split, the HBM footprint
    and the relative drift of total energy (kinetic + potential, FP64 reductions) and momentum."""
    n_plus = 256 ** 3 + 7_200_000
    n_total = n_plus + 256 ** 3
    log(f"generating the multistep workload (N={n_total}, PM 512^3) ...")
    t0 = time.perf_counter()
    w = synth.clustered(n_total=n_total, n_plus=n_plus, box_size=300.0, n_cells=512, n_halos=8000, seed=11)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s (N={w.n})")
    n = w.n
    pos0 = j.to_fixed(w.pos, w.box_size)
    steps = args.multisteps

    tuned = {}

    def params(prec):
        p = j.default_params(w.box_size, w.n_cells)
        p.r_s, p.r_cut, p.theta, p.softening, p.leaf_size = w.r_s, w.r_cut, w.theta, w.softening, w.leaf_size
        for k, v in tuned.items():
            setattr(p, k, v)
        p.precision = prec
        return p

    if args.autotune:
        # leaf and group size from the library's tuner, timed on the initial particles themselves
        t0 = time.perf_counter()
        res = autotune.autotune(w.pos, w.species, params(j.FP32), w.m_plus, w.m_minus_abs, k=1)
        tuned.update(leaf_size=res["leaf_size"], group_size=res["group_size"], fft_tile=res["fft_tile"])
        log(f"autotune: leaf {res['leaf_size']}, group {res['group_size']}, fft tile {res['fft_tile']} in "
            f"{time.perf_counter() - t0:.1f} s (sample {res['sample']}; leaf, group, ms, tree error: {res['table']}; "
            f"fft tile, ms: {res['fft_table']})")
    energy_ctx = j.ForceContext(params(j.FP64), w.m_plus, w.m_minus_abs)

    energy_ms = []

    def energy(dp, ds, dv, ctx):
        t = time.perf_counter()
        wpot = energy_ctx.potential_energy(n, dp, ds)
        energy_ms.append((time.perf_counter() - t) * 1e3)
        d = ctx.diagnostics(n, dv, ds)
        mom = d["momentum"].sum(axis=0)
        return wpot + d["kinetic"].sum(), mom, wpot

    out = {"metric": "TreePM KDK steps at N=4.08e7, PM 512^3, 1 GPU", "n_particles": n, "steps": steps,
           "pm_grid": 512, "box": 300.0, "n_halos": 8000, "forms": {},
           "leaf_size": params(j.FP32).leaf_size, "group_size": params(j.FP32).group_size,
           "fft_tile": params(j.FP32).fft_tile,
           "autotuned": bool(args.autotune)}
    dt = None
    for prec, label in ((j.FP32, "fp32"), (j.FP64, "fp64")):
        rt = np.float64 if prec == j.FP64 else np.float32
        free0, total = j.device_memory()
        ctx = j.ForceContext(params(prec), w.m_plus, w.m_minus_abs)
        dp, ds = j.DeviceArray.from_host(pos0), j.DeviceArray.from_host(w.species)
        dv, da = j.DeviceArray.from_host(np.zeros((n, 3), rt)), j.DeviceArray((n, 3), rt)
        ctx.compute_forces_device(n, dp, ds, da)
        free1, _ = j.device_memory()
        if dt is None:
            a = np.linalg.norm(da.download().astype(np.float64), axis=1)
            h = w.box_size / w.n_cells
            dt = float(np.sqrt(2.0 * h / np.percentile(a, 99)) / steps)  # p99 particle moves ~1 cell in total
        e0, p0, w0 = energy(dp, ds, dv, ctx)
        ctx.step_device(n, dp, dv, ds, da, dt, scheme=j.STEP_KDK, graph=False)  # one timed eager step
        split = ctx.stats()
        j.load_library().jtp_device_synchronize()
        t = time.perf_counter()
        for k in range(steps - 1):
            ctx.step_device(n, dp, dv, ds, da, dt, scheme=j.STEP_KDK, graph=True)
            if (k + 1) % 20 == 0:
                log(f"  {label}: step {k + 2}/{steps}")
        j.load_library().jtp_device_synchronize()
        ms = (time.perf_counter() - t) / max(1, steps - 1) * 1e3
        e1, p1, w1 = energy(dp, ds, dv, ctx)
        d = ctx.diagnostics(n, dv, ds)
        mscale = float(np.abs(d["momentum"]).sum())
        out["forms"][label] = {
            "ms_per_step": ms, "dt": dt, "first_step_split_ms": split,
            "hbm_footprint_bytes": int(free0 - free1) + 0,
            "energy_start": e0, "energy_end": e1, "potential_start": w0, "potential_end": w1,
            "kinetic_end": float(d["kinetic"].sum()),
            "rel_energy_drift": abs(e1 - e0) / abs(e0),
            "momentum_end": p1.tolist(), "rel_momentum_drift": float(np.abs(p1 - p0).max() / max(mscale, 1e-300)),
        }
        log(f"  {label}: {ms:.1f} ms/step, dE/E = {out['forms'][label]['rel_energy_drift']:.2e}")
        ctx.close()
        for arr in (dp, ds, dv, da):
            arr.free()
    out["fp32_over_fp64"] = out["forms"]["fp64"]["ms_per_step"] / out["forms"]["fp32"]["ms_per_step"]
    out["ms_potential_energy"] = min(energy_ms[1:]) if len(energy_ms) > 1 else None  # FP64, theta = 0 short range
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=22_000_000)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", default="clustered_22M",
                    choices=sorted(WORKLOADS) + ["ensemble_c3", "multistep_41M"])
    ap.add_argument("--multisteps", type=int, default=100, help="KDK steps of the multistep_41M workload")
    ap.add_argument("--ensemble-streams", type=int, default=4, help="streams of the ensemble_c3 workload")
    ap.add_argument("--autotune", action="store_true",
                    help="multistep_41M: pick leaf and group size with janus_treepm.autotune first")
    ap.add_argument("--no-c0", action="store_true", help="skip the full timing of the small configuration")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the comparison of both forms with the FP64 Ewald truth fixture")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's streaming copy kernel on MIB MiB, 5 times (HBM counter calibration)")
    ap.add_argument("--set", action="append", default=[], metavar="FIELD=VALUE",
                    help="override a jtp_params field (e.g. interlace=0, theta_cut=0.5, theta=0.4)")
    args = ap.parse_args()
    for kv in args.set:
        k, v = kv.split("=")
        OVERRIDES[k] = float(v) if k in ("theta", "theta_cut", "r_s", "r_cut", "softening") else int(v)
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the TreePM step is single-GPU; "
              "multi-GPU sharding is design-only) -- run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if j.device_count() <= 0:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        a, b = j.DeviceArray((nb,), np.uint8), j.DeviceArray((nb,), np.uint8)
        for _ in range(5):
            assert j.load_library().jtp_copy_d2d(b.ptr, a.ptr, nb) == 0
        print(json.dumps({"copy_calibration_bytes": nb, "launches": 5}))
        return

    if args.workload == "ensemble_c3":
        run_ensemble(args)
        return
    if args.workload == "multistep_41M":
        run_multistep(args)
        return

    t0 = time.perf_counter()
    cfg = dict(WORKLOADS[args.workload])
    if args.n != 22_000_000:
        cfg["n_total"] = args.n
    w = synth.clustered(**cfg)
    if args.workload == "clustered_344k":
        w.softening = synth.default_softening(250.0, 22_000_000)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s (N={w.n})")
    pos = j.to_fixed(w.pos, w.box_size)
    d_pos = j.DeviceArray.from_host(pos)
    d_sp = j.DeviceArray.from_host(w.species)

    t32, st32, _ = run_form(w, d_pos, d_sp, j.FP32, args.warmup, args.steps)
    ms32 = t32 / args.steps * 1e3
    value = w.n * args.steps / t32
    ph32 = phase_summary(st32)
    inter = int(np.mean([s["interactions"] for s in st32]))
    inter_mono = int(np.mean([s["interactions_mono"] for s in st32]))

    pdef = j.default_params(w.box_size, w.n_cells)
    for k, v in OVERRIDES.items():
        setattr(pdef, k, v)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-force-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"clustered two-species box N={w.n} ({cfg['n_halos']} NFW halos + uniform m+, "
                               f"void-filling m-), L={w.box_size:g}, PM {w.n_cells}^3, r_s=1.25 cells, "
                               f"r_cut=4.5 r_s, theta={w.theta}, leaf {w.leaf_size}, m+=1, |m-|=19; "
                               f"FP32 fixed-point form",
                   "n_particles": w.n, "pm_grid": w.n_cells, "theta": w.theta, "r_s": w.r_s, "r_cut": w.r_cut,
                   "softening": w.softening, "leaf_size": w.leaf_size, "group_size": pdef.group_size,
                   "theta_cut": pdef.theta_cut, "parallelism": "single GPU", "overrides": OVERRIDES},
    }
    walk_ms = ph32["ms_walk"]
    achieved = FLOPS_PER_INTERACTION * inter / (walk_ms * 1e-3) / 1e12
    traffic, traffic_src = (None, None)
    if not OVERRIDES and args.n == 22_000_000:
        traffic, traffic_src = measured_traffic(args.workload)
    out["roofline"] = {"bound": "valu", "kernel": "k_walk (group walk, FP32)", "achieved": achieved,
                       "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                       "traffic": traffic, "traffic_source": traffic_src, "interactions": inter,
                       "interactions_mono": inter_mono, "flops_per_interaction": FLOPS_PER_INTERACTION,
                       "walk_ms": walk_ms}
    out["phases_ms"] = ph32
    # PM kernels alone: in the step the tree build shares the GPU with them (side stream), so their
    # HBM rates are taken from PM-only evaluations (tree disabled) of the same particles.
    saved = dict(OVERRIDES)
    OVERRIDES["tree_enabled"] = 0
    _, st_pm, _ = run_form(w, d_pos, d_sp, j.FP32, 2, 5)
    OVERRIDES.clear()
    OVERRIDES.update(saved)
    ph_pm = phase_summary(st_pm)
    G3 = w.n_cells ** 3
    grids = 2 if OVERRIDES.get("interlace", 1) else 1  # interlaced PM runs every grid phase twice
    hbm = {
        "deposit": (w.n * 13 + grids * 4 * G3) / (ph_pm["ms_deposit"] * 1e-3) / 1e9,
        # 48 G^3 per solve: 3 passes x (read + write) x 4 B x G^3 x 2 directions, the algorithmic
        # figure of the measurement plan (the kernels move 40 G^3: 5 passes over the half spectrum)
        "fft": (grids * 48 * G3) / (ph_pm["ms_fft"] * 1e-3) / 1e9,
        "interp": (w.n * 25 + grids * 4 * G3) / (ph_pm["ms_interp"] * 1e-3) / 1e9,
    }
    out["phase_hbm_gbs"] = {k: {"achieved": v, "frac": v / PEAK_HBM_GBS, "ms": ph_pm["ms_" + k]}
                            for k, v in hbm.items()}

    if not args.no_fp64:
        t64, st64, _ = run_form(w, d_pos, d_sp, j.FP64, 1, args.fp64_steps)
        out["fp64"] = {"value": w.n * args.fp64_steps / t64, "ms_per_step": t64 / args.fp64_steps * 1e3,
                       "steps": args.fp64_steps, "phases_ms": phase_summary(st64)}
        out["fp32_over_fp64"] = out["fp64"]["ms_per_step"] / ms32

    if not args.no_accuracy:
        log("accuracy vs Ewald truth ...")
        out["accuracy_vs_ewald"] = accuracy(w, args.workload, d_pos, d_sp, pos)
    del pos

    if not args.no_c0:
        log("config 0 (2 x 64^3, PM 64^3) ...")
        out["config0"] = config0()

    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
