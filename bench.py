"""Throughput of the pseudospectral vorticity step on one MI355X.

Workload (the 4096^2 forced-turbulence case the headline metric is quoted on):
  2-D periodic [0, 2 pi]^2, N = 4096, nu = 1e-4 (Re = 1e4), del^8
  hyperviscosity nu4 = 1e-16 applied as an integrating factor, linear drag
  0.1, white-in-time forcing on the shell 4 <= |k| <= 6 at injection rate
  1e-3 (counter SplitMix64, seed 42), adaptive dt at CFL 0.5, integrating-
  factor RK4 stages in FP32 (FP64 measured beside it).  Initial vorticity:
  random phases with E(k) ~ k^4 exp(-(k/10)^2), unit energy, seed 42
  (synthetic data).

A step is one full RK4 step (4 x {row kernel, column kernel} + the forcing
draw) over the whole grid, state resident in HBM.  The timed region runs K
steps the way ns_spectral_advance_n does, one captured step graph replayed K
times (eager launches with --no-graph), host clock, the solver's stream
synchronized on both sides.  The same K steps are then launched eagerly, and
K more with a HIP event around every stage kernel on the solver's stream:
the per-kernel averages give the roofline of the dominant kernel.  Prints one
JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# oracle steps of the 4096^2 workload timed for the CPU baseline
CPU_STEPS = 8
METRIC = "Gpoint-steps/s (4096², FP32 vs FP64), % HBM peak, closure speedup@equal error"


def workload(ns, n, fp32):
    s = ns.Spectral(n, n, 2 * np.pi, 2 * np.pi, 1e4)
    s.init(taylor_green=False)
    kmax = np.sqrt(2.0) * n / 2.0  # hyperviscosity.c:68-73 (corner of the stored modes)
    s.hv_set(enabled=1, coefficient=1e-16 * kmax ** 8, order=8, cutoff_ratio=0.0)
    s.set_precision(fp32)
    s.set_scheme(True)
    s.set_drag(0.1)
    s.set_forcing(4.0, 6.0, 1e-3, seed=42)
    s.d.cfl_factor = 0.5
    s.random_phase(42, k_peak=10.0, energy=1.0)
    s.sync_device()
    return s


def kernel_bytes(n, fp32):
    """HBM bytes per launch of the two stage kernels, two ways.

    Pipeline model (the roofline's algorithmic figure): a stage of the
    reference's pipeline moves 16 half-spectrum fields of n (n/2 + 1) complex
    values -- K1 (read w_s, write 4 derivative fields: 5), K2 (read 4, write
    J: 5), K3 (read J, w_n, w_s, read and write the accumulator, write w_s+1:
    6).  k_row is K2 (5 fields), k_col fuses K3 and K1 (11 fields).

    Own bytes: what these kernels must move in their layouts -- mixed fields
    hold only the kept kx columns, compact fields only the kept modes."""
    nkx = n // 2 + 1
    kxk = (2 * nkx) // 3 + 1            # kept kx columns
    kyk = 2 * ((2 * n) // 6) + 1        # kept ky rows
    esz = 8 if fp32 else 16             # complex element
    M = kxk * n * esz                   # one mixed field [kx][y]
    Cc = kxk * kyk * esz                # one compact spectral field
    Hf = n * nkx * esz                  # one half-spectrum field of the model
    own = {"k_row": 5 * M,              # 4 mixed in, J out
           "k_col": 5 * M + 2.75 * Cc}  # J in, 4 mixed out, state/acc traffic averaged
    model = {"k_row": 5 * Hf, "k_col": 11 * Hf}
    return own, model


PROBE_BYTES = 1 << 30


# Clock ramp: the GPU leaves idle clocks over the first ~70 ms of load (about
# 100 steps at 4096^2), and steps timed inside that window run several
# percent slow.  Whatever --warmup asks for, at least this many untimed steps
# run before the timed region; the extra ones are reported as "ramp_steps".
RAMP_MIN_STEPS = 100


def time_gpu(ns, n, fp32, steps, warmup, probe=False, compat=0, graph=True, ramp_min=0):
    L = ns.load()
    s = workload(ns, n, fp32)
    ctx = s.ctx
    probe_gbs = None
    if probe:
        ms = C.c_double()
        s.check(L.nsd_bandwidth_probe(ctx, PROBE_BYTES, 5, C.byref(ms)))
        probe_gbs = 2 * PROBE_BYTES / (ms.value * 1e-3) / 1e9
    ramp = max(0, ramp_min - warmup)
    if ramp:
        s.check(L.nsd_advance_graph(ctx, ramp, 1) if graph else L.nsd_advance(ctx, ramp, 1))
    if warmup:
        s.check(L.nsd_advance_graph(ctx, warmup, 1) if graph else L.nsd_advance(ctx, warmup, 1))
    s.check(L.nsd_synchronize(ctx))
    # timed region: K steps through the path of ns_spectral_advance_n (one
    # captured step graph replayed K times; eager launches with graph=False),
    # the solver's stream synchronized on both sides, host clock (nothing else
    # on the stream)
    t0 = time.perf_counter()
    s.check(L.nsd_advance_graph(ctx, steps, 1) if graph else L.nsd_advance(ctx, steps, 1))
    s.check(L.nsd_synchronize(ctx))
    wall = time.perf_counter() - t0
    # the same number of steps launched eagerly, kernel by kernel (host clock)
    eager_wall = wall
    if graph:
        t0 = time.perf_counter()
        s.check(L.nsd_advance(ctx, steps, 1))
        s.check(L.nsd_synchronize(ctx))
        eager_wall = time.perf_counter() - t0
    # the next K steps again, with HIP events on the solver's stream around
    # every stage kernel: the per-kernel durations of the roofline (the
    # events themselves add a few microseconds per kernel boundary)
    L.nsd_profile(ctx, 1)
    s.check(L.nsd_advance(ctx, steps, 1))
    s.check(L.nsd_synchronize(ctx))
    L.nsd_profile(ctx, 0)
    row_ms, col_ms, step_ms = C.c_double(), C.c_double(), C.c_double()
    nst = C.c_int()
    L.nsd_profile_read(ctx, C.byref(row_ms), C.byref(col_ms), C.byref(step_ms), C.byref(nst))
    st = s.sync_host()
    e, z = s.device_diagnostics()
    res = dict(ms_per_step=1e3 * wall / steps, events_ms_per_step=step_ms.value,
               launch="graph replay" if graph else "eager",
               eager_ms_per_step=1e3 * eager_wall / steps, row_ms=row_ms.value,
               col_ms=col_ms.value, t=s.d.t, dt=s.d.dt, energy=e, enstrophy=z, status=st,
               probe_gbs=probe_gbs, compat_ms_per_step=None, ramp_steps=ramp)
    if compat:
        # the reference entry point: every call uploads omega_hat and downloads
        # omega_hat, psi_hat, u_hat, v_hat, u, v, omega over PCIe (host arrays)
        s.advance_rk4()  # first call allocates the FP64 staging buffers
        t0 = time.perf_counter()
        for _ in range(compat):
            s.advance_rk4()
        res["compat_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / compat
    s.close()
    return res


def ensemble_workload(members=256, seed=2024):
    """Sizes uniform over {128, 256, 512, 1024}, nu log-uniform in [1e-5, 1e-3],
    one seed per member; forcing, drag and del^8 hyperviscosity as the 4096^2
    workload, FP32 integrating-factor stages."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 512, 1024], size=members)
    nu = np.exp(rng.uniform(np.log(1e-5), np.log(1e-3), size=members))
    seeds = np.arange(1, members + 1)
    return sizes, 1.0 / nu, seeds


def time_ensemble(ns, steps, warmup, members=256):
    sizes, Re, seeds = ensemble_workload(members)
    e = ns.Ensemble(sizes, Re, seeds, forcing=(4.0, 6.0, 1e-3), precision=ns.NS_PRECISION_FP32,
                    scheme=ns.NS_SCHEME_IF_RK4, cfl_factor=0.5, drag=0.1, nu4=1e-16,
                    k_peak=10.0, energy=1.0)
    e.advance(warmup)
    e.synchronize()
    t0 = time.perf_counter()
    e.advance(steps)
    e.synchronize()
    wall = time.perf_counter() - t0
    t, E, Z, st = e.diagnostics()
    e.close()
    pts = float(np.sum(np.asarray(sizes, dtype=np.float64) ** 2))
    return {"workload": f"{members} members, sizes uniform over 128..1024, nu log-uniform "
                        "[1e-5, 1e-3], FP32 IF-RK4, forcing/drag/del^8 as the 4096^2 workload",
            "members_steps_per_s": round(members * steps / wall, 1),
            "gpoint_steps_per_s": round(pts * steps / wall / 1e9, 3),
            "ms_per_ensemble_step": round(1e3 * wall / steps, 4),
            "members_per_size": {int(n): int(np.sum(sizes == n)) for n in (128, 256, 512, 1024)},
            "all_status_ok": bool(np.all(st == 0))}


def closure_workload(ns, seeds=64, snapshots=500, stride=10, epochs=10, members=16, t_end=0.5,
                     train_only=False):
    """Learned closure: fine 2048^2 forced runs (one per seed) sampled every
    `stride` steps into (coarse vorticity, subgrid tendency) pairs at 256^2
    (64 x 500 = 32,000 pairs, 16.8 GB in HBM), Adam training of the bf16 MFMA
    network, then 256^2 ensembles with and without it against the filtered
    2048^2 ensemble from the same initial fields (same physics as the 4096^2
    workload).  A closure that fails the enstrophy check hands back the fine
    solve (`result`, `result_speedup_vs_fine`)."""
    import importlib
    cl = importlib.import_module(ns.__name__ + ".closure")
    net = cl.Net(256, max_batch=max(16, members), seed=1)
    t0 = time.perf_counter()
    net.reserve(seeds * snapshots)
    cl.generate(net, 2048, list(range(1, seeds + 1)), snapshots, stride)
    t_data = time.perf_counter() - t0
    print(f"closure: {seeds * snapshots} training pairs in {t_data:.1f} s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    l0, l1, lv = net.train(epochs, batch=16, lr=2e-3)
    t_train = time.perf_counter() - t0
    print(f"closure: {epochs} epochs in {t_train:.1f} s", file=sys.stderr, flush=True)
    if train_only:  # counter runs of the network kernels: no evaluation
        net.close()
        return {"workload": f"2048^2 -> 256^2, {seeds} seeds x {snapshots} snapshots every "
                            f"{stride} steps, {epochs} epochs; training only",
                "loss": [round(l0, 4), round(l1, 4)], "validation_loss": round(lv, 4),
                "data_s": round(t_data, 1), "train_s": round(t_train, 1)}
    r = cl.evaluate(net, 2048, list(range(1001, 1001 + members)), t_end)
    net.close()
    return {"workload": f"2048^2 -> 256^2, {seeds} seeds x {snapshots} snapshots every "
                        f"{stride} steps, {epochs} epochs; evaluation: {members}-member "
                        f"ensembles to t = {r['t_end']:.3f}",
            "speedup_vs_fine": round(r["closure"]["speedup_vs_fine"], 2),
            "spectrum_mean_rel_err": round(r["closure"]["spectrum_mean_rel_err"], 4),
            "spectrum_max_rel_err": round(r["closure"]["spectrum_max_rel_err"], 4),
            "spectrum_max_shell": r["closure"]["spectrum_max_shell"],
            "spectrum_max_shell_level": float(f'{r["closure"]["spectrum_max_shell_level"]:.3g}'),
            "enstrophy_rel_err": round(r["closure"]["enstrophy_rel_err"], 4),
            "result_speedup_vs_fine": round(r["result_speedup_vs_fine"], 2),
            "plain_coarse": {"spectrum_mean_rel_err": round(r["plain"]["spectrum_mean_rel_err"], 4),
                             "spectrum_max_rel_err": round(r["plain"]["spectrum_max_rel_err"], 4),
                             "enstrophy_rel_err": round(r["plain"]["enstrophy_rel_err"], 4),
                             "speedup_vs_fine": round(r["plain"]["speedup_vs_fine"], 2)},
            "result": r["result"], "loss": [round(l0, 4), round(l1, 4)],
            "validation_loss": round(lv, 4), "data_s": round(t_data, 1),
            "train_s": round(t_train, 1)}


def cpu_baseline(n):
    """Oracle (reference algorithm, project FFT, OpenMP) on this host at FP64:
    CPU_STEPS steps of the same forced integrating-factor workload (the figure
    that sits beside the GPU value, about 10 s of CPU work on 16 cores at
    4096^2), then one step each of the reference's own
    classic RK4 (spectral_time_integration.c:34-106, with its redundant
    transforms) and of the same RK4 with only the transforms it needs, on the
    same field."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(n, n, 2 * np.pi, 2 * np.pi, 1e4)
    o.init_grids()
    kmax = np.sqrt(2.0) * n / 2.0
    o.hv_set(enabled=1, coefficient=1e-16 * kmax ** 8, order=8, cutoff_ratio=0.0)
    o.d.drag = 0.1
    o.d.cfl_factor = 0.5
    o.init_random_phase(42, 10.0, 1.0)
    f = O.OracleForcing(o.k2, n // 2 + 1, n, 4.0, 6.0, energy_target=1.0, seed=42)
    t0 = time.perf_counter()
    for _ in range(CPU_STEPS):
        o.d.dt = o.cfl()
        o.advance_if(o.project_forcing(f.generate(o.d.dt), 1e-3))
    t = time.perf_counter() - t0
    out = dict(value=CPU_STEPS * n * n / t / 1e9, unit="Gpoint-steps/s", cores=threads,
               kind="port",
               sample=f"{CPU_STEPS} RK4 steps of the {n}^2 forced integrating-factor workload "
                      f"with the CFL step and forcing draws, FP64, oracle with project FFT "
                      f"(not FFTW), {t:.2f} s")
    # the reference's step itself (no forcing, drag or integrating factor)
    o.d.drag = 0.0
    o.compute_derivatives()
    for key, minimal in (("reference_rk4_step", False), ("minimal_rk4_step", True)):
        c0 = o.d.fft_count
        t0 = time.perf_counter()
        o.advance_rk4(minimal=minimal)
        t = time.perf_counter() - t0
        out[key] = {"value": round(n * n / t / 1e9, 5), "unit": "Gpoint-steps/s",
                    "seconds": round(t, 2), "ffts": int(o.d.fft_count - c0)}
    o.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1, help="GPUs (this build measures one)")
    ap.add_argument("--steps", type=int, default=50, help="timed steps (K)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ramp", type=int, default=RAMP_MIN_STEPS,
                    help="untimed steps before the timed region, warm-up included (clock ramp)")
    ap.add_argument("--n", type=int, default=4096, help="grid size N of the N^2 workload")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline (eight oracle steps and two reference steps, "
                         "about 25 s)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 step")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip training and evaluating the learned closure (about 100 s)")
    ap.add_argument("--closure-snapshots", type=int, default=500,
                    help="training snapshots per fine run (64 runs) of the closure leg")
    ap.add_argument("--closure-train-only", action="store_true",
                    help="closure leg without its evaluation, for counter runs of the network "
                         "kernels")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 256-member mixed-size ensemble")
    ap.add_argument("--no-compat", action="store_true",
                    help="skip timing the PCIe-bound reference entry point ns_spectral_advance_rk4")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches only (no captured-graph warm-up or replay), for tools "
                         "that cannot follow graph launches")
    ap.add_argument("--probe", action="store_true",
                    help="also time a plain 1 GiB device copy (k_copy) for counter calibration")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as G
    ns = G._import_package()
    n = a.n
    r32 = time_gpu(ns, n, True, a.steps, a.warmup, a.probe, compat=0 if a.no_compat else 3,
                   graph=not a.no_graph, ramp_min=a.ramp)
    r64 = None if a.no_fp64 else time_gpu(ns, n, False, max(a.steps // 2, 1), max(a.warmup // 2, 1),
                                          graph=not a.no_graph)
    value = n * n / (r32["ms_per_step"] * 1e-3) / 1e9

    own, model = kernel_bytes(n, True)
    # dominant kernel: the longer of the two kernel-only event brackets
    kms = {"k_row": r32["row_ms"], "k_col": r32["col_ms"]}
    dom = max(kms, key=kms.get)
    dom_s = kms[dom] * 1e-3
    achieved = model[dom] / dom_s / 1e9
    achieved_own = own[dom] / dom_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_fp32_{n}.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        if f"{dom}<float" in tj.get("kernel", ""):
            traffic = tj.get("hbm_bytes_per_launch")
    step_bytes = 4 * (own["k_row"] + own["k_col"])
    # the pipeline model per step: 16 half-spectrum fields per stage, 64 per
    # step, 8 bytes per FP32 complex value
    pipeline_model = 64 * n * (n // 2 + 1) * 8
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpoint-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ramp_steps": r32["ramp_steps"],
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"forced 2-D turbulence {n}^2, FP32 stages, integrating-factor "
                               "RK4, del^8 nu4=1e-16, drag 0.1, white forcing 4<=|k|<=6 rate "
                               "1e-3 seed 42, adaptive dt CFL 0.5, nu=1e-4",
                   "grid": n, "precision": "fp32"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": model[dom],
                     "algorithmic_model": "pipeline model, 16 half-spectrum fields per stage: "
                                          "k_row 5, k_col 11",
                     "kernel_us": round(kms[dom] * 1e3, 2),
                     "own_bytes_per_launch": own[dom],
                     "achieved_own_bytes": round(achieved_own, 1),
                     "frac_own_bytes": round(achieved_own / HBM_PEAK_GBS, 4)},
        "step": {"launch": r32["launch"],
                 "row_ms": round(r32["row_ms"], 4), "col_ms": round(r32["col_ms"], 4),
                 "eager_ms_per_step": round(r32["eager_ms_per_step"], 4),
                 "events_ms_per_step": round(r32["events_ms_per_step"], 4),
                 "algorithmic_bytes_per_step": step_bytes,
                 "step_hbm_frac": round(step_bytes / (r32["ms_per_step"] * 1e-3) / 1e9
                                        / HBM_PEAK_GBS, 4),
                 "pipeline_model_hbm_frac": round(pipeline_model / (r32["ms_per_step"] * 1e-3)
                                                  / 1e9 / HBM_PEAK_GBS, 4),
                 "t": r32["t"], "dt": r32["dt"], "energy": r32["energy"], "status": r32["status"]},
    }
    if r32["compat_ms_per_step"]:
        out["compat_ms_per_step"] = round(r32["compat_ms_per_step"], 2)
    if r32["probe_gbs"]:
        out["copy_probe_gbs"] = round(r32["probe_gbs"], 1)
    if r64:
        out["fp64"] = {"value": round(n * n / (r64["ms_per_step"] * 1e-3) / 1e9, 3),
                       "ms_per_step": round(r64["ms_per_step"], 4), "launch": r64["launch"],
                       "eager_ms_per_step": round(r64["eager_ms_per_step"], 4),
                       "row_ms": round(r64["row_ms"], 4), "col_ms": round(r64["col_ms"], 4)}
    if not a.no_ensemble:
        out["ensemble"] = time_ensemble(ns, max(a.steps // 2, 5), 3)
    if not a.no_closure:
        out["closure"] = closure_workload(ns, snapshots=a.closure_snapshots,
                                          train_only=a.closure_train_only)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
