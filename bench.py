#!/usr/bin/env python3
"""Benchmark: Mcell-updates/s per RK stage of the RHS + Runge-Kutta path.

Workload (one GPU): supersonic NS Taylor-Green vortex, 256^3 periodic
interior (262^3 with ghosts), Ma 1.25, Re 1600, SLAU ORDER 6 with HR-SLAU2 and
hybrid TVD (threshold 0.4), central-6 viscous terms, TVD-RK3, CFL 0.03
(synthetic initial field, no I/O).  One "step" is one full RK3 step (three
stages).  `value` is the FP64 plain form; the FP32-bulk form of the same
workload is measured beside it.

A serving leg (the Kelvin-Helmholtz request stream through the batch engine)
reports p50/p99 solve latency beside throughput.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu] [--serve-requests N]
"""
import argparse
import ctypes
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def _opcounts():
    rows = [json.loads(l) for l in (ROOT / "ouxsbli_amd" / "flops" / "opcount.json").read_text().splitlines() if l]
    return {r["config"]: r for r in rows}


def kernel_model(spec, ops):
    """Compulsory HBM bytes and algorithmic FP ops per launch of each kernel of
    the stage pipeline.  Bytes count every array element a kernel must read or
    write once: primitives and the residual accumulator in the form's element
    type (8 B FP64, 4 B FP32-bulk), the state in FP64.  Ops come from
    ouxsbli_amd/flops/opcount.json."""
    padded = spec.nx * spec.ny * spec.nz
    cells = spec.interior_cells
    e = 8 if spec.precision == "fp64" else 4
    np_ = 5  # rho, u, v, w, p: the sweeps form T and mu from rho and p
    face = ops["conv_face_smooth"] + 2 + ops["visc_face"] + 4 * ops["cross_deriv"]
    m = {
        # stale primitives only (after a host upload)
        "prim": {"bytes": padded * (40 + np_ * e), "ops": padded * 20},
        "sweep_x": {"bytes": cells * (np_ * e + e + 5 * e), "ops": cells * (face + 10)},
        # reads the primitives and the sensor; reads and writes the residual
        "sweep_y": {"bytes": cells * ((np_ + 1) * e + 10 * e), "ops": cells * (face + 15)},
        # reads the primitives, the sensor and the residual; RK3 average of
        # the state traffic (stage 1: Qin read, Qout write; stages 2-3: also
        # Qout read); writes the primitives of the new state
        "sweep_z_rk": {"bytes": cells * ((np_ + 1) * e + 5 * e + 40 + 40 + 40 * 2 / 3 + np_ * e),
                       "ops": cells * (face + 15 + 4 * 5 + 20)},
        "ghost_fill": {"bytes": (padded - cells) * 80, "ops": 0},
    }
    return m


# Measured HBM bytes per launch of the FP64 kernels at 256^3 (rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of this script, scaled by what the same
# counters read for a device copy of known size), tagged with the SHA-256 of
# the kernel sources they were measured on: a kernel change makes them stale,
# and the bench line then reports traffic null.
TRAFFIC_FILE = ROOT / "ouxsbli_amd" / "flops" / "traffic.json"
KERNEL_SOURCES = ("ox_kernels.hpp", "ox_physics.hpp", "ox_stage_impl.hpp", "ox_types.hpp", "ox_core.hpp",
                  "ox_fp64.hip")


def kernel_sources_sha():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update((ROOT / "ouxsbli_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()


def measured_traffic(n):
    """{kernel: HBM bytes per launch} measured on the current kernel sources
    at this grid, or None."""
    if not TRAFFIC_FILE.exists():
        return None
    t = json.loads(TRAFFIC_FILE.read_text())
    if t.get("sources_sha256") != kernel_sources_sha() or t.get("n") != n:
        return None
    return t["bytes_per_launch"]


def host_cpu():
    """CPU model, the CPUs this process may run on and their physical cores."""
    model, phys = "unknown", set()
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    try:
        cur = {}
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines() + [""]:
            if not line.strip():
                if cur.get("processor") is not None and int(cur["processor"]) in cpus:
                    phys.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                    model = cur.get("model name", model)
                cur = {}
                continue
            k, _, v = line.partition(":")
            cur[k.strip()] = v.strip()
    except OSError:
        pass
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(phys) or len(cpus)}


def _oracle_rate(spec, nsteps, repeats):
    """Timing protocol: 1 warm-up step, then `repeats` timings of
    `nsteps` steps with a steady clock; the median rate (Mcell-updates/s per
    RK stage).  Setup and I/O excluded."""
    from ouxsbli_amd import cases
    import oracle_lib
    o = oracle_lib.Oracle(spec)
    o.set_state(cases.initial_state(spec))
    o.step(1)
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.step(nsteps)
        rates.append(spec.interior_cells * spec.stages * nsteps / (time.perf_counter() - t0) / 1e6)
    rates.sort()
    return rates[len(rates) // 2], rates


def cpu_baseline(n_small=64, n_large=128):
    """The FP64 OpenMP oracle (the parity build) on the host cores, on bounded
    samples: the headline scheme at 64^3 (10 steps, median of 5) and 128^3
    (2 steps, median of 3), and the 64^3 KEEP4 NS TGV in full (RK4,
    100 steps).  OpenMP threads are bound to cores (OMP_PROC_BIND=close,
    OMP_PLACES=cores) before the oracle library loads; OMP_NUM_THREADS keeps
    the value the environment sets, else one thread per physical core."""
    cpu = host_cpu()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_NUM_THREADS", str(cpu["physical_cores"]))
    sys.path.insert(0, str(ROOT / "tests"))
    from ouxsbli_amd import cases
    t0 = time.perf_counter()
    v64, r64 = _oracle_rate(cases.supersonic_tgv_spec(n_small, "fp64"), 10, 5)
    v128, r128 = _oracle_rate(cases.supersonic_tgv_spec(n_large, "fp64"), 2, 3)
    # the 64^3 KEEP4 NS TGV in full
    import oracle_lib
    c1 = cases.tgv_keep4_spec(64)
    o = oracle_lib.Oracle(c1)
    o.set_state(cases.initial_state(c1))
    tc = time.perf_counter()
    o.step(100)
    t_c1 = time.perf_counter() - tc
    threads = int(os.environ["OMP_NUM_THREADS"])
    return {"value": round(v128, 2), "unit": "Mcell-updates/s per RK stage", "cores": threads, "kind": "port",
            "sample": f"oracle (oracle/liboxoracle.so, FP64 OpenMP, g++ -O3 -ffp-contract=off) on the headline scheme "
                      f"(SLAU6 HR-SLAU2 hybrid NS6 RK3) at {n_large}^3: 1 warm-up + 3 x 2 timed steps, median; "
                      f"{time.perf_counter() - t0:.1f} s in all",
            "cpu_model": cpu["model"], "physical_cores": cpu["physical_cores"], "logical_cpus": cpu["logical_cpus"],
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "headline_scheme_64": {"value": round(v64, 2), "steps": 10, "repeats": [round(r, 2) for r in r64]},
            "headline_scheme_128": {"value": round(v128, 2), "steps": 2, "repeats": [round(r, 2) for r in r128]},
            "tgv64_keep4_full": {"value": round(c1.interior_cells * c1.stages * 100 / t_c1 / 1e6, 2),
                             "seconds": round(t_c1, 2),
                             "workload": "NS TGV 64^3 KEEP4 central-4 viscous RK4 Ma 0.1, 100 steps"}}


def les_leg(n, steps):
    """LES of decaying isotropic turbulence: seeded random-phase solenoidal
    field, E(k) ~ k^4 exp(-2 (k/4)^2), Mt 0.3, uniform rho and T; Hybrid
    KEEP/SLAU (Ducros threshold 0.4), SMS SGS model, central-6 viscous terms,
    RK4.  The FP32-bulk form (the configured one) is timed beside the FP64
    plain form, which stands in for the oracle at this size (it is
    bit-identical to the oracle on the same scheme at 24^3 in the tests)."""
    from ouxsbli_amd import cases
    spec32 = cases.dhit_spec(n, "fp32_bulk")
    spec64 = cases.dhit_spec(n, "fp64")
    out = {"workload": f"LES decaying isotropic turbulence {n}^3 periodic, E(k)~k^4 exp(-2(k/4)^2), Mt 0.3, "
                       "Hybrid KEEP/SLAU ORDER 6 (threshold 0.4), SMS SGS, central-6 viscous, RK4; initial field "
                       "synthesised on the device from its spectral modes",
           "steps": steps}
    res = {}
    for name, spec in (("fp32_bulk", spec32), ("fp64", spec64)):
        r = run_form_state(spec, None, steps, 1)
        res[name] = r
        cs = spec.interior_cells * spec.stages * steps
        out[name] = {"Mcell_updates_per_s_stage": round(cs / r["wall_s"] / 1e6, 1),
                     "ms_per_step": round(r["wall_s"] * 1e3 / steps, 3),
                     "initial_state_s": round(r["ic_s"], 3) if r["ic_s"] is not None else None}
    k32, k64 = res["fp32_bulk"]["diag"]["ke"], res["fp64"]["diag"]["ke"]
    out["fp32_bulk"]["ke_rel_err_vs_fp64"] = abs(k32 - k64) / abs(k64)
    out["fp32_bulk"]["speedup_vs_fp64"] = round(out["fp32_bulk"]["Mcell_updates_per_s_stage"] /
                                                out["fp64"]["Mcell_updates_per_s_stage"], 3)
    out["state_ok"] = bool(res["fp32_bulk"]["ok"] and res["fp64"]["ok"])
    return out


def run_form_state(spec, Q, steps, warmup):
    from ouxsbli_amd.solver import Solver
    with Solver(spec) as s:
        if Q is None:
            t_ic = time.perf_counter()
            s.init_state()
            ic_s = time.perf_counter() - t_ic
        else:
            s.set_state(Q)
            ic_s = None
        if warmup:
            s.step(warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        diag = s.diagnostics()
        ok, _, _ = s.check_state()
    return {"wall_s": wall, "diag": diag, "ok": ok, "ic_s": ic_s}


OS_BYTES_PER_CELL_STAGE = (64.0 + 96.0 + 96.0) / 3.0  # 2-D RK3: stage 1 reads Q, writes Q1; stages 2, 3 read two


def batched_os_leg(ncase, steps, unbatched_steps, cpu_seconds):
    """Batched ensemble of 2-D oblique-shock solves: 64 cases, interior grids
    nx-2 ~ U{128..512}, ny-2 ~ U{64..256}, Mach 2, deflection ~ U(4, 12) deg
    (seeded), Rankine-Hugoniot state on the top boundary, Roe + MUSCL (ORDER 6,
    hybrid limiter), RK3, FP32 bulk, `steps` steps each, all resident and
    advanced together: one persistent launch per RK stage over the tiles of
    all cases, replayed from a HIP graph.  Timed beside: the FP64 plain form
    of the same batch, and the FP32 cases advanced one by one (one launch per
    stage per case), in both forms."""
    import numpy as np
    from ouxsbli_amd import cases2d as c2
    from ouxsbli_amd.batch import Batch
    from ouxsbli_amd.solver2d import Solver2D
    out = {"workload": f"{ncase} batched 2-D oblique-shock reflections, interior 128..512 x 64..256, M 2, "
                       "theta U(4,12) deg, Roe + MUSCL ORDER 6 hybrid limiter, RK3, "
                       f"{steps} steps each, one persistent launch per stage with HIP graph replay",
           "cases": ncase, "steps": steps}
    res = {}
    for prec in ("fp32_bulk", "fp64"):
        specs = c2.oblique_shock_specs(ncase, precision=prec)
        states = [c2.os_state(s) for s in specs]
        cs = sum(s.interior_cells for s in specs) * 3 * steps
        with Batch(hbm_budget=32 << 30, max_cases=ncase, use_graph=True, autotune=False) as b:
            ids = [b.admit(s, Q, steps) for s, Q in zip(specs, states)]
            t0 = time.perf_counter()
            while b.run(steps)[1]:
                pass
            wall = time.perf_counter() - t0
            st = b.stats()
            final = {cid: b.retire(cid) for cid in ids}
        res[prec] = [final[cid] for cid in ids]
        dev = st["device_ms"] * 1e-3
        out[prec] = {"Mcell_updates_per_s_stage": round(cs / wall / 1e6, 1),
                     "device_busy_Mcell_updates_per_s_stage": round(cs / dev / 1e6, 1),
                     "wall_s": round(wall, 3), "graph_replays": st["graph_replays"],
                     "avg_stage_launch_ms": round(st["device_ms"] / (3 * steps), 4),
                     "alg_GBs": round(cs / dev * OS_BYTES_PER_CELL_STAGE / 1e9, 1)}
        out["cell_stages"] = cs
    out["fp32_bulk"]["speedup_vs_fp64"] = round(out["fp32_bulk"]["Mcell_updates_per_s_stage"] /
                                                out["fp64"]["Mcell_updates_per_s_stage"], 3)
    # the fused stage kernel against both roofs: compulsory FP64 state bytes and
    # the algorithmic operation count of smooth faces (ouxsbli_amd/flops/opcount.json;
    # limited faces cost more, so this understates the work done)
    ops = _opcounts()["roe6_hybrid_2d_rk3"]["per_cell_stage_smooth"]
    rate = out["fp32_bulk"]["device_busy_Mcell_updates_per_s_stage"] * 1e6
    out["fp32_bulk"]["roofline"] = {
        "kernel": "k2d_stage<float,6,ROE,*>", "bound": "hbm", "unit": "GB/s",
        "achieved": round(rate * OS_BYTES_PER_CELL_STAGE / 1e9, 1), "peak": HBM_PEAK_GBS,
        "frac": round(rate * OS_BYTES_PER_CELL_STAGE / 1e9 / HBM_PEAK_GBS, 4),
        "ops_per_cell_stage": ops, "achieved_TFLOPs": round(rate * ops / 1e12, 2), "peak_TFLOPs": FP32_PEAK_TFLOPS,
        "flop_frac": round(rate * ops / 1e12 / FP32_PEAK_TFLOPS, 4)}
    # accuracy of the FP32-bulk ensemble against the FP64 one, per case
    errs = []
    for a, b in zip(res["fp64"], res["fp32_bulk"]):
        d = a[:, 0, :] - b[:, 0, :]
        errs.append(float(np.sqrt(np.mean(d * d) / np.mean(a[:, 0, :] ** 2))))
    out["fp32_bulk"]["rho_rel_l2_vs_fp64_max"] = max(errs)
    out["fp32_bulk"]["rho_rel_l2_vs_fp64_median"] = float(np.median(errs))
    # the same cases one at a time through the plain single-case kernels (one
    # launch per stage per case), in both forms
    for prec in ("fp32_bulk", "fp64"):
        specs = c2.oblique_shock_specs(ncase, precision=prec)
        t_dev = 0.0
        cs1 = 0
        for s in specs:
            with Solver2D(s) as h:
                h.set_state(c2.os_state(s))
                h.step(unbatched_steps)
                t_dev += h.last_step_ms() * 1e-3
            cs1 += s.interior_cells * 3 * unbatched_steps
        out[f"{prec}_one_case_at_a_time"] = {"device_Mcell_updates_per_s_stage": round(cs1 / t_dev / 1e6, 1),
                                             "steps": unbatched_steps}
    out["fp32_bulk"]["speedup_vs_fp64_one_case_at_a_time"] = round(
        out["fp32_bulk"]["device_busy_Mcell_updates_per_s_stage"] /
        out["fp64_one_case_at_a_time"]["device_Mcell_updates_per_s_stage"], 3)
    if cpu_seconds > 0:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib
        s = c2.os_spec(258, 130, scheme="Roe", cfl=c2.CFL_REF)
        o = oracle_lib.Oracle2D(s)
        o.set_state(c2.os_state(s))
        o.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < cpu_seconds:
            o.step(10)
            n += 10
        el = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        out["cpu_oracle"] = {"value": round(s.interior_cells * 3 * n / el / 1e6, 1),
                             "unit": "Mcell-updates/s per RK stage", "cores": cores, "kind": "port",
                             "sample": f"2-D oracle (FP64 OpenMP), one 256 x 128 Roe case, {n} steps ({el:.1f} s)"}
    return out


SERVE_BUDGET = 64 << 30
SERVE_SLOTS = 16


def _serve_once(n_requests, precision, seed, batch=None, nsteps=200):
    from ouxsbli_amd import serve
    reqs = serve.khi_stream(n_requests, rate_hz=20.0, nsteps=nsteps, precision=precision, with_states=False,
                            seed=seed)
    rep = serve.serve(reqs, hbm_budget=SERVE_BUDGET, max_cases=SERVE_SLOTS, keep_results=False, pretune=False,
                      batch=batch)
    s = rep.summary()
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}


def serving_leg(n_requests, fp64_value):
    """Kelvin-Helmholtz serving stream through the batch engine: seeded
    Poisson arrivals at 20 req/s, interior grids 64^3..256^3 (multiples of
    16), KEEP or SLAU at ORDER 4 or 6, Euler, RK3, 200 steps each; FIFO
    admission by HBM footprint (64 GiB budget) and 16 slots, each request's
    initial state built on the device when it is admitted.

    - cold: a fresh engine; segment lengths are tuned when a shape is first
      admitted, inside that request's latency;
    - warm: an engine that served an earlier stream (another seed, 2 steps per
      request) keeps its tuned shapes, as a running server does; shapes it
      has not seen are tuned on admission as in the cold run;
    - the same stream in the FP64 plain form on a warm engine, for the
      FP32-bulk/FP64 ratio on identical requests."""
    from ouxsbli_amd.batch import Batch
    seed = 20261207
    out = {"workload": "3-D Kelvin-Helmholtz stream: double tanh shear layer d=L/28, Ma 0.5, 1e-3 U seeded noise, "
                       f"{n_requests} requests, Poisson 20 req/s, grids 64^3-256^3, KEEP/SLAU ORDER 4/6, Euler RK3, "
                       "200 steps each, FP32-bulk, initial states built on the device at admission, batched per "
                       "stage with HIP graph replay",
           "hbm_budget_GiB": SERVE_BUDGET >> 30, "slots": SERVE_SLOTS}
    cold = _serve_once(n_requests, "fp32_bulk", seed)
    out.update(cold)
    out["tuning"] = "cold: each new shape tuned at admission, inside latency"
    warm = {}
    for prec in ("fp32_bulk", "fp64"):
        with Batch(SERVE_BUDGET, max_cases=SERVE_SLOTS, use_graph=True, autotune=True) as b:
            _serve_once(n_requests, prec, seed + 1000, batch=b, nsteps=2)
            warm[prec] = _serve_once(n_requests, prec, seed, batch=b)
    out["warm"] = warm["fp32_bulk"]
    out["fp64_served_warm"] = warm["fp64"]
    b32 = warm["fp32_bulk"]["device_busy_Mcell_updates_per_s_stage"]
    b64 = warm["fp64"]["device_busy_Mcell_updates_per_s_stage"]
    out["fp32_vs_fp64_same_requests_device_busy"] = round(b32 / b64, 3) if b32 and b64 else None
    out["fp32_vs_fp64_same_requests_wall"] = round(warm["fp32_bulk"]["Mcell_updates_per_s_stage"] /
                                                   warm["fp64"]["Mcell_updates_per_s_stage"], 3)
    busy = cold["device_busy_Mcell_updates_per_s_stage"]
    out["device_busy_vs_fp64_plain_supersonic_tgv"] = round(busy / fp64_value, 3) if busy else None
    return out


def les_serving_leg(n_requests, nsteps=50):
    """LES request stream through the batch engine: decaying isotropic
    turbulence (Hybrid KEEP/SLAU ORDER 6, SMS model, central-6 viscous, RK4),
    seeded Poisson arrivals at 4 req/s, interior grids 64^3..192^3
    (multiples of 32), `nsteps` steps each, each request's turbulent field
    synthesised on the device at admission from its own spectral seed.  Warm
    engines (an earlier stream of the same shapes, 2 steps each), FP32-bulk
    and FP64 on the same requests."""
    from ouxsbli_amd import serve
    from ouxsbli_amd.batch import Batch
    seed = 20261208
    out = {"workload": f"LES isotropic-turbulence stream: {n_requests} requests, Poisson 4 req/s, grids "
                       f"64^3-192^3, Hybrid KEEP/SLAU ORDER 6, SMS SGS, central-6 viscous, RK4, {nsteps} steps "
                       "each, fields synthesised on the device at admission, batched per stage (SGS viscosity "
                       "included) with HIP graph replay",
           "hbm_budget_GiB": SERVE_BUDGET >> 30, "slots": SERVE_SLOTS}
    res = {}
    for prec in ("fp32_bulk", "fp64"):
        with Batch(SERVE_BUDGET, max_cases=SERVE_SLOTS, use_graph=True, autotune=True) as b:
            serve.serve(serve.dhit_stream(n_requests, seed=seed + 1000, nsteps=2, precision=prec),
                        hbm_budget=SERVE_BUDGET, max_cases=SERVE_SLOTS, keep_results=False, batch=b)
            rep = serve.serve(serve.dhit_stream(n_requests, seed=seed, nsteps=nsteps, precision=prec),
                              hbm_budget=SERVE_BUDGET, max_cases=SERVE_SLOTS, keep_results=False, batch=b)
        res[prec] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.summary().items()}
    out["fp32_bulk"] = res["fp32_bulk"]
    out["fp64"] = res["fp64"]
    b32 = res["fp32_bulk"]["device_busy_Mcell_updates_per_s_stage"]
    b64 = res["fp64"]["device_busy_Mcell_updates_per_s_stage"]
    out["fp32_vs_fp64_same_requests_device_busy"] = round(b32 / b64, 3) if b32 and b64 else None
    return out


def run_form(spec, steps, warmup, profile):
    from ouxsbli_amd import cases
    from ouxsbli_amd.solver import Solver
    Q = cases.tgv_state(spec)
    with Solver(spec) as s:
        s.set_state(Q)
        lseg = s.autotune()  # the library's segment-length tuner on this grid (the state is unchanged)
        if warmup:
            s.step(warmup)
        s.set_profiling(profile)
        s.synchronize()
        t0 = time.perf_counter()
        dev_ms = s.step(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        kt = s.kernel_times() if profile else {}
        diag = s.diagnostics()
        ok, _, _ = s.check_state()
    return {"wall_s": wall, "dev_ms": dev_ms, "kernels": kt, "diag": diag, "ok": ok, "lseg": lseg}


def history_leg(n, nsteps, every):
    """The 256^3 supersonic TGV's kinetic-energy and enstrophy histories: the FP64
    plain and the FP32-bulk forms from the same initial field, diagnostics
    every `every` steps over `nsteps` steps (the FP64 GPU form stands in for
    the oracle at this size; it equals the oracle bit for bit on the same
    scheme at the sizes the oracle runs, tests/test_gpu_seams.py).  Also the
    sustained rate of each form: device time of all steps, diagnostics
    excluded."""
    from ouxsbli_amd import cases
    from ouxsbli_amd.solver import Solver
    out = {"workload": f"{n}^3 supersonic NS TGV (SLAU6 HR-SLAU2 hybrid NS6 RK3), {nsteps} steps",
           "every": every}
    hist = {}
    for prec in ("fp64", "fp32_bulk"):
        spec = cases.supersonic_tgv_spec(n, prec)
        rows = []
        dev_ms = 0.0
        t0 = time.perf_counter()
        with Solver(spec) as s:
            s.init_state()
            s.autotune()
            d = s.diagnostics()
            rows.append((0, d["ke"], d["enstrophy"]))
            for k in range(every, nsteps + 1, every):
                dev_ms += s.step(every)
                d = s.diagnostics()
                rows.append((k, d["ke"], d["enstrophy"]))
            ok, _, _ = s.check_state()
        wall = time.perf_counter() - t0
        hist[prec] = rows
        cs = spec.interior_cells * spec.stages * nsteps
        out[prec] = {"sustained_Mcell_updates_per_s_stage": round(cs / (dev_ms * 1e-3) / 1e6, 1),
                     "device_s": round(dev_ms * 1e-3, 3), "wall_s_with_diagnostics": round(wall, 3),
                     "state_ok": bool(ok)}
    ke_err = max(abs(a[1] - b[1]) / abs(a[1]) for a, b in zip(hist["fp64"], hist["fp32_bulk"]))
    en_err = max(abs(a[2] - b[2]) / abs(a[2]) for a, b in zip(hist["fp64"], hist["fp32_bulk"]))
    out["fp32_bulk"]["ke_rel_err_max"] = ke_err
    out["fp32_bulk"]["enstrophy_rel_err_max"] = en_err
    out["fp32_bulk"]["sustained_speedup_vs_fp64"] = round(out["fp32_bulk"]["sustained_Mcell_updates_per_s_stage"] /
                                                          out["fp64"]["sustained_Mcell_updates_per_s_stage"], 3)
    out["history"] = {"step": [r[0] for r in hist["fp64"]],
                      "ke_fp64": [r[1] for r in hist["fp64"]], "ke_fp32_bulk": [r[1] for r in hist["fp32_bulk"]],
                      "enstrophy_fp64": [r[2] for r in hist["fp64"]],
                      "enstrophy_fp32_bulk": [r[2] for r in hist["fp32_bulk"]]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-me4", dest="me4", action="store_false",
                    help="skip the ME4Base-viscous variant of the workload")
    ap.add_argument("--history-steps", type=int, default=500,
                    help="steps of the configuration-2 KE/enstrophy history in both forms (0: skip it)")
    ap.add_argument("--les-n", type=int, default=384,
                    help="interior size of the LES decaying-turbulence leg (0: skip it)")
    ap.add_argument("--les-steps", type=int, default=30)
    ap.add_argument("--os-cases", type=int, default=64,
                    help="cases of the batched 2-D oblique-shock leg (0: skip it)")
    ap.add_argument("--os-steps", type=int, default=2000)
    ap.add_argument("--os-only", action="store_true",
                    help="run only the batched oblique-shock leg and print its JSON (for profiling)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the Kelvin-Helmholtz serving stream (0: skip the serving leg)")
    ap.add_argument("--les-serve-requests", type=int, default=12,
                    help="requests of the LES serving stream (0: skip it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the library holds each case on one GPU "
              "and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.os_only:
        print(json.dumps(batched_os_leg(args.os_cases, args.os_steps, 100, 0.0)), flush=True)
        return

    from ouxsbli_amd import cases
    spec = cases.supersonic_tgv_spec(args.n, "fp64")
    spec32 = cases.supersonic_tgv_spec(args.n, "fp32_bulk")
    ops = _opcounts()["slau6_hrslau2_hybrid_ns6_rk3"]

    r64 = run_form(spec, args.steps, args.warmup, True)
    r32 = run_form(spec32, args.steps, args.warmup, True)

    cells = spec.interior_cells
    cell_stages = cells * spec.stages * args.steps
    value = cell_stages / r64["wall_s"] / 1e6
    value32 = cell_stages / r32["wall_s"] / 1e6
    ke64, ke32 = r64["diag"]["ke"], r32["diag"]["ke"]

    # The roofline of the dominant kernel, as the algorithmic bytes of the
    # work one launch does (106.7 B per cell-stage of RK3: the compulsory FP64
    # state reads and writes, primitives and residual recomputed on chip, per
    # cell of the launch) over its average launch time, measured live with HIP
    # events on the library's stream.
    B_alg = 320.0 / 3.0
    model = kernel_model(spec, ops)
    kt = r64["kernels"]
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    alg_bytes = B_alg * cells
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.n)
    kernels = {}
    for k, (tot, n) in kt.items():
        a = tot / n
        kernels[k] = {"avg_ms": round(a, 4), "share": round(tot / sum(v[0] for v in kt.values()), 4),
                      "compulsory_GBs": round(model[k]["bytes"] / (a * 1e-3) / 1e9, 1),
                      "measured_traffic_GBs": round(traffic[k] / (a * 1e-3) / 1e9, 1) if traffic and k in traffic
                      else None,
                      "TFLOPs": round(model[k]["ops"] / (a * 1e-3) / 1e12, 2)}
    F_alg = ops["per_cell_stage_smooth"]
    stage_rate = value * 1e6
    hbm_rate = HBM_PEAK_GBS * 1e9 / B_alg
    fp64_rate = FP64_PEAK_TFLOPS * 1e12 / F_alg
    bound_rate = min(hbm_rate, fp64_rate)
    stage_ms = r64["wall_s"] * 1e3 / (args.steps * spec.stages)

    line = {
        "metric": "Mcell-updates/s per RK stage, FP32-bulk vs FP64 plain; p50/p99 solve latency",
        "value": round(value, 1),
        "unit": "Mcell-updates/s per RK stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["wall_s"] * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"supersonic NS TGV {args.n}^3 periodic, Ma 1.25, Re 1600, SLAU ORDER 6 HR-SLAU2 "
                               "hybrid-TVD (threshold 0.4), central-6 viscous, TVD-RK3, CFL 0.03",
                   "interior": [args.n] * 3, "extents": [spec.nx, spec.ny, spec.nz], "stages_per_step": 3,
                   "form": "FP64 plain (bit-identical to the CPU oracle)",
                   "segment_lengths_yz": list(r64["lseg"])},
        "fp32_bulk": {"value": round(value32, 1), "ms_per_step": round(r32["wall_s"] * 1e3 / args.steps, 3),
                      "speedup_vs_fp64": round(value32 / value, 3),
                      "ke_rel_err_vs_fp64": abs(ke32 - ke64) / abs(ke64)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get(dom) if traffic else None,
                     "alg_bytes_per_cell_stage": round(B_alg, 2), "cells_per_launch": cells,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(avg_ms, 4),
                     "kernel_compulsory_bytes_per_launch": model[dom]["bytes"],
                     "kernel_compulsory_GBs": round(model[dom]["bytes"] / (avg_ms * 1e-3) / 1e9, 1),
                     "traffic_source": "ouxsbli_amd/flops/traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                       "tagged with the kernel sources' SHA-256)" if traffic else
                                       "no measurement on the current kernel sources"},
        "stage_roofline": {"stage_ms": round(stage_ms, 4),
                           "F_alg_per_cell_stage": F_alg, "B_alg_per_cell_stage": round(B_alg, 1),
                           "hbm_frac": round(stage_rate / hbm_rate, 4),
                           "fp64_frac": round(stage_rate / fp64_rate, 4),
                           "binding": "fp64" if fp64_rate < hbm_rate else "hbm",
                           "bound_Mcell_stage_s": round(bound_rate / 1e6, 1),
                           "frac": round(stage_rate / bound_rate, 4)},
        "kernels": kernels,
        "state_ok": bool(r64["ok"] and r32["ok"]),
    }
    if args.me4:
        spec_me4 = cases.supersonic_tgv_spec(args.n, "fp64")
        spec_me4.visc_stencil = "ME4Base"
        r_me4 = run_form(spec_me4, args.steps, args.warmup, False)
        v_me4 = spec_me4.interior_cells * spec_me4.stages * args.steps / r_me4["wall_s"] / 1e6
        line["viscous_me4base"] = {"value": round(v_me4, 1), "ms_per_step": round(r_me4["wall_s"] * 1e3 / args.steps, 3),
                                   "form": "FP64 plain, the same workload with the ME4Base viscous stencil"}
    if args.history_steps > 0:
        line["accuracy_history"] = history_leg(args.n, args.history_steps, 10)
    if args.les_n > 0:
        line["les_dhit"] = les_leg(args.les_n, args.les_steps)
    if args.os_cases > 0:
        line["batched_oblique_shock"] = batched_os_leg(args.os_cases, args.os_steps, 100,
                                                       0.0 if args.no_cpu else 5.0)
    if args.serve_requests > 0:
        line["serving"] = serving_leg(args.serve_requests, value)
    if args.les_serve_requests > 0:
        line["les_serving"] = les_serving_leg(args.les_serve_requests)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
