"""Benchmark: particle-steps/s of the MFiX DEM hot path on one MI355X.

Workload (default): FP64 CFD-DEM with a frozen gas field -- 1,000,000
polydisperse spheres (log-normal d, median 1 mm, sigma 0.25, clipped to
[0.5, 2] mm), Hertz contact (E = 1e7 Pa, nu = 0.3), e = 0.9, mu = 0.3,
gas grid 64 x 128 x 64 with 1.2 m/s upflow plus a seeded 5 % sinusoidal
perturbation, Gidaspow drag, GARG_2012 trilinear deposition, dt = 1e-5 s.
A "step" is one DES_TIME_STEP over all particles (walls, pairs, drag,
integration, fluid-cell binning, deposition, EP_G; a neighbour-list rebuild
every 25 steps).  Particle state is resident in HBM before timing starts.

Prints ONE JSON line.  `value` = FP64 particle-steps/s; the mixed-precision
run of the same workload is reported beside it.  `roofline` covers the
dominant kernel (the fused particle step) timed with HIP events;
`cpu_baseline` times the CPU restatement (OpenMP) on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mfix-cfd-dem_amd"))

import numpy as np  # noqa: E402

from mfixdem import abi, workloads  # noqa: E402
from mfixdem.system import DemSystem  # noqa: E402

METRIC = "particle-steps/s at 1M spheres FP64 vs mixed; p50/p99 sweep-request latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_step_kernel(n, pairs, dep_nodes=0):
    """Bytes one fused step launch must move: per particle read pos, vel,
    omega, radius and write pos, vel, omega (152 B); per half-list pair the
    neighbour index (4 B) and the tangential history read + write (48 B);
    with the GARG_2012 deposition fused in, per gas node its int64 solids
    sum read and written once (16 B)."""
    return 152.0 * n + 52.0 * pairs + 16.0 * dep_nodes


def make_system(p, parts, gas, tuning=None):
    s = DemSystem(p)
    s.set_particles(parts)
    s.set_gas(gas)
    if tuning is not None:
        s.set_tuning(*tuning)
    s.init()
    return s


def tune_single(p, parts, gas, log):
    """Result-invariant tuning of the neighbour-row kernel for the workload
    (every candidate gives bit-identical results)."""
    from mfixdem import autotune

    def make():
        s = DemSystem(p)
        s.set_particles(parts)
        s.set_gas(gas)
        return s
    return autotune.autotune(make, cands=[(1.0, abi.ROWS_THREAD), (1.0, abi.ROWS_WAVE128),
                                          (1.0, abi.ROWS_WAVE192)], log=log)


def time_steps(s, nn0, steps):
    s.synchronize()
    t0 = time.perf_counter()
    s.run(nn0, steps)
    s.synchronize()
    return time.perf_counter() - t0


def timed_repeats(s, nn0, steps, repeats, k):
    """`repeats` timed windows of exactly `steps` steps.  Window r starts at
    NN = nn0 + r*L with L the smallest multiple of the rebuild interval k
    that is >= steps (untimed steps fill the gaps), so every window holds
    the same number of neighbour rebuilds.  Returns (seconds per window,
    next NN)."""
    period = -(-steps // k) * k
    out, nn = [], nn0
    for r in range(max(1, repeats)):
        if r:
            s.run(nn, period - steps)
            nn += period - steps
        out.append(time_steps(s, nn, steps))
        nn += steps
    return out, nn


def kernel_times(s, nn0, steps):
    """Per-kernel-group HIP event times over `steps` eager steps."""
    s.set_debug(abi.DEBUG_TIME_KERNELS)
    s.lib.mfixdem_reset_kernel_times(s.h)
    s.run(nn0, steps)
    arr = (C.c_double * abi.TIME_GROUPS)()
    calls = (C.c_int64 * abi.TIME_GROUPS)()
    s.lib.mfixdem_get_kernel_times(s.h, arr, calls, abi.TIME_GROUPS)
    s.set_debug(0)
    return [arr[i] / max(calls[i], 1) for i in range(abi.TIME_GROUPS)], list(calls)


def kinetic_energy(parts, vel, om):
    """Translational + rotational kinetic energy of the particle set (J)."""
    r = parts.radius
    m = (4.0 / 3.0) * np.pi * r ** 3 * parts.ro_sol
    inertia = 0.4 * m * r * r
    return float(0.5 * np.sum(m * np.sum(vel * vel, axis=0)) +
                 0.5 * np.sum(inertia * np.sum(om * om, axis=0)))


def bed_height(parts, pos, frac=0.95):
    """Height (y) below which `frac` of the solid volume lies."""
    vol = (4.0 / 3.0) * np.pi * parts.radius ** 3
    order = np.argsort(pos[1])
    c = np.cumsum(vol[order])
    return float(pos[1][order][np.searchsorted(c, frac * c[-1])])


def force_error(p, parts, gas):
    """Max |F_mixed - F_fp64| / max |F_fp64| over all particles for the
    first step from the same state (FC as the integrator sees it)."""
    out = []
    for prec in (abi.FP64, abi.MIXED):
        q = abi.Params.from_buffer_copy(p)
        q.precision = prec
        s = DemSystem(q)
        s.set_particles(parts)
        s.set_gas(gas)
        s.set_debug(abi.DEBUG_STORE_FORCES)
        s.init()
        s.step(1)
        out.append(s.get_forces()[0])
        s.close()
    return float(np.max(np.abs(out[1] - out[0])) / np.max(np.abs(out[0])))


def skin_bench(p, parts, gas, vel, tuning, steps, value_ref, log):
    """The opt-in skin tuner on the same bed: FACTOR_RLM candidates, each
    with the longest rebuild interval that keeps every contact for speeds
    up to twice the largest one seen in the FP64 run, timed from the same
    initial state.  Not the reference's schedule (results differ in the
    last bits), so reported beside the headline value, not as it."""
    from mfixdem import autotune
    vmax = float(np.sqrt((vel ** 2).sum(0)).max())
    r_min = float(parts.radius.min())

    def make(q):
        s = DemSystem(q)
        s.set_particles(parts)
        s.set_gas(gas)
        s.set_tuning(*tuning)
        return s
    (f, n), timings = autotune.autotune_skin(make, p, r_min, vmax, factors=(1.2, 1.35, 1.5),
                                             steps=100, log=log)
    q = abi.Params.from_buffer_copy(p)
    q.factor_rlm = f
    q.neighbor_search_n = n
    s = make(q)
    s.init()
    s.run(1, 50)
    el = time_steps(s, 51, steps)
    s.close()
    v = parts.n * steps / el
    log(f"skin: FACTOR_RLM {f}, NEIGHBOR_SEARCH_N {n}: {v:.4e} particle-steps/s "
        f"(reference schedule {value_ref:.4e})")
    return {"factor_rlm": f, "neighbor_search_n": n, "v_max_probe": vmax, "safety": 2.0,
            "r_min": r_min, "particle_steps_per_s": v, "candidates_s": timings,
            "reference_schedule_particle_steps_per_s": value_ref}


def ensemble_bench(nsys, steps, log):
    """Parameter-sweep ensemble: nsys independent settling beds of uneven size (N log-uniform
    in [2k, 200k], d = 1 mm), KN / EN / MEW swept, DTSOLID = TCOLL/50 each, in
    one ensemble handle; the result-invariant tuning is chosen by the
    autotuner on a 32-member sample first."""
    import functools
    from mfixdem import autotune
    from mfixdem.ensemble import Ensemble
    member = functools.partial(workloads.lattice_member, precision=abi.FP64)
    t0 = time.perf_counter()
    ms = workloads.sweep_ensemble(nsys=nsys, seed=11, member=member)
    log(f"ensemble: {nsys} systems, {sum(p.max_pip for p, _ in ms)} particles built in "
        f"{time.perf_counter() - t0:.1f} s")

    def make(sub):
        e = Ensemble([p for p, _ in sub])
        for s, (_, parts) in enumerate(sub):
            e.set_particles(s, parts)
        return e

    sample = ms[:: max(1, nsys // 32)]
    best, timings = autotune.autotune(
        lambda: make(sample),
        cands=autotune.candidates((1.0, 1.25), (abi.ROWS_THREAD, abi.ROWS_WAVE128,
                                                abi.ROWS_WAVE192)), log=log)
    e = make(ms)
    e.set_tuning(*best)
    e.init()
    e.run(1, 25)
    e.synchronize()
    t0 = time.perf_counter()
    e.run(26, steps)
    e.synchronize()
    el = time.perf_counter() - t0
    npart = sum(p.max_pip for p, _ in ms)
    out = {"systems": nsys, "particles": npart, "steps": steps,
           "particle_steps_per_s": npart * steps / el, "ms_per_step": 1e3 * el / steps,
           "device_bytes": e.device_bytes(),
           "tuning": {"cell_scale": best[0], "row_kernel": best[1],
                      "candidates_s": timings},
           "data": "synthetic (seeded jittered-lattice packings, swept KN/EN/MEW)"}
    e.close()
    log(f"ensemble: {out['particle_steps_per_s']:.4e} particle-steps/s, tuning {best}")
    return out


def serving_bench(n_req, log, rate=20.0, workers=8, budget=64 << 30):
    """The sweep-request stream: Poisson arrivals at `rate` per second of
    DEM runs with N log-uniform in [5k, 500k] and steps uniform in [200,
    5000], mixed precision, KN/EN/MEW swept; each request is built when it
    is admitted (lazily), admitted by HBM footprint within `budget`, and
    served alongside the running ones.  The first n_req requests of the
    seeded stream (2,000 for the full stream)."""
    from mfixdem import serving
    reqs = serving.poisson_requests(n_req, rate=rate, make_member=workloads.lattice_member,
                                    nmin=5000, nmax=500_000, smin=200, smax=5000, seed=21,
                                    lazy=True)
    model = serving.FootprintModel(lambda n: workloads.lattice_member(n, 1e3, 0.9, 0.3, 1))
    for r in reqs:
        r.footprint = model(r.n)
    work = sum(r.work for r in reqs)
    span = reqs[-1].arrival
    srv = serving.SweepServer(budget, workers=workers)
    done = srv.serve(reqs)
    rep = serving.report(done, srv.makespan)
    rep.update({"arrival_rate_per_s": rate, "offered_particle_steps_per_s": work / span,
                "last_arrival_s": span, "drain_after_last_arrival_s": srv.makespan - span,
                "hbm_budget_bytes": budget, "workers": workers,
                "max_concurrent": srv.max_running, "max_bytes_in_use": srv.max_in_use,
                "workload": f"{n_req} sweep requests: N log-uniform [5k, 500k], steps U[200, "
                            f"5000], mixed precision, KN/EN/MEW swept, Poisson arrivals at "
                            f"{rate}/s, built on admission"})
    log(f"serving: {n_req} requests, p50 {rep['p50_latency_s']:.3f} s, p99 "
        f"{rep['p99_latency_s']:.3f} s, served {rep['particle_steps_per_s']:.3e} of offered "
        f"{rep['offered_particle_steps_per_s']:.3e} particle-steps/s, max {srv.max_running} "
        f"running")
    return rep


def two_way_bench(s, reps=20):
    """Gas side of the drag (DRAG_GS_GAS0: PARTICLES_IN_CELL lists, F_GP per
    particle, DRAG_AM/DRAG_BM node gather in the reference's order, F_GDS)
    on the current state: host wall time per call, stream synchronised."""
    abi.check(s.lib.mfixdem_drag_gs_gas(s.h), "drag_gs_gas")   # buffers, node volumes
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        abi.check(s.lib.mfixdem_drag_gs_gas(s.h), "drag_gs_gas")
    s.synchronize()
    us = (time.perf_counter() - t0) / reps * 1e6
    f_gds, _, _ = s.drag_gs_gas()
    return {"us_per_call": us, "calls": reps, "f_gds_max": float(f_gds.max()),
            "routine": "DRAG_GS_GAS0 (F_GDS, DRAG_AM, DRAG_BM), reference summation order"}


def cpu_model():
    """CPU model and logical CPU count of this host (lscpu, else /proc/cpuinfo)."""
    import subprocess
    model, cpus = None, os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                model = v.strip()
            elif k.strip() == "CPU(s)":
                cpus = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    return model, cpus


def cpu_baseline(p, parts, gas, cycles_all=3, cycles_one=1):
    """Times the OpenMP CPU restatement (oracle/liboracle_omp.so) on the same
    workload over whole rebuild cycles: steps NN = K+1 .. 2K are K steps
    with exactly one neighbour rebuild (at NN = 2K), as in the GPU run.
    All threads (OMP_NUM_THREADS, else every CPU): median over `cycles_all`
    consecutive cycles; 1 thread: `cycles_one` cycle(s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import OracleSystem

    K = p.neighbor_search_n
    model, cpus = cpu_model()

    def timed(threads, cycles):
        o = OracleSystem(p, omp=True)
        o.lib.orc_set_omp_threads(threads)
        used = int(o.lib.orc_omp_threads())
        o.set_particles(parts)
        o.set_gas(gas)
        o.init()
        o.run(1, K)                      # NN = 1 .. K, untimed
        per_cycle = []
        for c in range(cycles):
            t0 = time.perf_counter()
            o.run((c + 1) * K + 1, K)    # NN = (c+1)K+1 .. (c+2)K, one rebuild
            per_cycle.append(time.perf_counter() - t0)
        o.close()
        return used, per_cycle

    n_all = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (cpus or 1)
    used_all, cyc_all = timed(n_all, cycles_all)
    used_one, cyc_one = timed(1, cycles_one)
    med = float(np.median(cyc_all))
    v_all = p.max_pip * K / med
    v_one = p.max_pip * K / float(np.median(cyc_one))
    return {"value": v_all, "unit": "particle-steps/s", "cores": used_all, "kind": "port",
            "sample": f"{p.max_pip}-sphere workload, OpenMP CPU restatement, {used_all} threads: "
                      f"median of {cycles_all} whole {K}-step cycles (each with one neighbour "
                      f"rebuild), {sum(cyc_all):.1f} s in total",
            "cpu_model": model, "host_cpus": cpus,
            "cycle_s": cyc_all,
            "single_thread": {"value": v_one, "cores": used_one, "cycle_s": cyc_one,
                              "sample": f"{cycles_one} whole {K}-step cycle(s) with one "
                                        f"rebuild each, 1 thread"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--cpu-cycles", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-skin", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--ens-systems", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one DEM system per GPU on a single GPU; "
              f"--gpus {args.gpus} is not supported (multi-GPU sharding is design-only)",
              file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    p, parts, gas = workloads.fluidised_bed(n=args.n)
    log(f"workload built in {time.perf_counter() - t0:.1f} s: n={parts.n}")

    tuning, tune_s = tune_single(p, parts, gas, log)
    # FP64
    s = make_system(p, parts, gas, tuning)
    s.run(1, args.warmup)
    reps, nn = timed_repeats(s, 1 + args.warmup, args.steps, args.repeats,
                             p.neighbor_search_n)
    el = float(np.median(reps))
    st = s.status()
    pairs = st.n_pairs
    value = parts.n * args.steps / el
    log(f"fp64: {args.steps} steps, median of {len(reps)} repeats {el:.4f} s "
        f"({', '.join(f'{r:.4f}' for r in reps)}) -> {value:.4e} particle-steps/s, "
        f"pairs/particle {pairs / parts.n:.2f}")
    kt, calls = kernel_times(s, nn, 50)
    t_step_ms = kt[abi.TIME_STEP]
    # the bench bed is GARG_2012, two-way coupled: the step kernel deposits
    dep_nodes = workloads.n_cells(p) if (p.coupled and p.interp_scheme == abi.INTERP_GARG_2012) else 0
    bytes_step = algorithmic_bytes_step_kernel(parts.n, pairs, dep_nodes)
    achieved = bytes_step / (t_step_ms * 1e-3) / 1e9
    log(f"fused step kernel {t_step_ms * 1e3:.1f} us, rebuild {kt[abi.TIME_REBUILD] * 1e3:.1f} us, "
        f"gas {kt[abi.TIME_GAS] * 1e3:.1f} us")
    nsteps_total = args.warmup + len(reps) * args.steps + 50
    pos64, vel64, om64, _ = s.get_particles()
    two_way = two_way_bench(s)
    log(f"two-way coupling: DRAG_GS_GAS0 {two_way['us_per_call']:.1f} us per call")
    s.close()

    # HBM bytes per k_step launch from the committed rocprofv3 counter runs
    # (FETCH_SIZE, WRITE_SIZE; a profiler run of its own), with the date they
    # were measured and whether they describe the kernel sources built here
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_step_fluidised_1M.json")
    if os.path.exists(tpath) and args.n == 1_000_000:
        from mfixdem import traffic as traffic_mod
        tj = json.load(open(tpath))
        traffic = tj.get("bytes_per_launch")
        traffic_src = {"file": "profiles/traffic_k_step_fluidised_1M.json",
                       "measured_utc": tj.get("measured_utc"),
                       "same_kernel_sources": tj.get("kernel_source_sha256_16") ==
                       traffic_mod.kernel_source_digest()}

    mixed = None
    if not args.no_mixed:
        pm = abi.Params.from_buffer_copy(p)
        pm.precision = abi.MIXED
        m = make_system(pm, parts, gas, tuning)
        m.run(1, args.warmup)
        repm, nnm = timed_repeats(m, 1 + args.warmup, args.steps, args.repeats,
                                  pm.neighbor_search_n)
        elm = float(np.median(repm))
        ktm, _ = kernel_times(m, nnm, 50)
        posm, velm, omm, _ = m.get_particles()
        m.close()
        ke64, kem = kinetic_energy(parts, vel64, om64), kinetic_energy(parts, velm, omm)
        ke0 = kinetic_energy(parts, parts.vel, parts.omega)
        h64, hm = bed_height(parts, pos64), bed_height(parts, posm)
        mixed = {"value": parts.n * args.steps / elm, "ms_per_step": 1e3 * elm / args.steps,
                 "step_kernel_us": ktm[abi.TIME_STEP] * 1e3,
                 "steps_compared": nsteps_total,
                 "max_rel_force_error": force_error(p, parts, gas),
                 "ke_fp64": ke64, "ke_mixed": kem, "ke_initial": ke0,
                 "ke_rel_diff": abs(kem - ke64) / ke64,
                 "bed_height_95_fp64_m": h64, "bed_height_95_mixed_m": hm,
                 "bed_height_rel_diff": abs(hm - h64) / h64}
        log(f"mixed: {mixed['value']:.4e} particle-steps/s, force err "
            f"{mixed['max_rel_force_error']:.2e}, KE diff {mixed['ke_rel_diff']:.2e}, "
            f"bed height {h64:.5f} / {hm:.5f} m")

    skin = None if args.no_skin else skin_bench(p, parts, gas, vel64, tuning, args.steps,
                                                value, log)
    ens = None if args.no_ensemble else ensemble_bench(args.ens_systems, args.ens_steps, log)
    srv = None if args.no_serving else serving_bench(args.serve_requests, log)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, parts, gas, cycles_all=args.cpu_cycles)
        log(f"cpu baseline: {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads, "
            f"{cpu['single_thread']['value']:.4e} on 1 thread ({cpu['cpu_model']})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "repeats_s": reps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded RSA packing, seeded frozen gas field)",
        "config": {"workload": "fluidised bed, FP64 CFD-DEM with frozen gas (Hertz, "
                               "Gidaspow, GARG_2012)",
                   "n_particles": parts.n, "gas_grid": [p.imax, p.jmax, p.kmax],
                   "des_grid": [p.desgridsearch_imax, p.desgridsearch_jmax,
                                p.desgridsearch_kmax],
                   "dt": p.dtsolid, "neighbor_search_n": p.neighbor_search_n,
                   "pairs_per_particle": pairs / parts.n,
                   "tuning": {"cell_scale": tuning[0], "row_kernel": tuning[1],
                              "candidates_s": tune_s}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_step (fused walls+pairs+drag+integrate+locate)",
                     "kernel_us": t_step_ms * 1e3, "algorithmic_bytes": bytes_step},
        "cpu_baseline": cpu,
        "mixed": mixed,
        "skin": skin,
        "ensemble": ens,
        "serving": srv,
        "kernel_groups_us": {"step": kt[0] * 1e3, "rebuild": kt[1] * 1e3, "gas": kt[2] * 1e3},
        "two_way": two_way,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
