"""Benchmark of the MI355X all-pairs gravity path (libabie_hip.so).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras]

A "step" is one acceleration evaluation of the force-kernel workload: 64
massive bodies (star + 63, m log-U[1e-9, 1e-3]) plus 1,048,576 test
particles (a ~ U[1, 100] au), FP64, inputs resident in HBM.  value =
algorithmic interactions (N_m (N_m - 1) + N_t N_m = 67,112,896 per step) per
second over the K timed steps.  The kernel's mean launch time
(roofline.achieved) comes from HIP events on the launch stream around at
least 20 ms of back-to-back launches, whatever K is; its HBM traffic from
the committed rocprofv3 counter summary (profiles/force_c1a_traffic.json).

Extras (reported beside the headline, not part of it): the 65,536-body
Plummer force case (and the fast FP64 mode's difference from the exact mode
on every sink), the mixed-precision and exact force modes, the drop-in GR15
on the 1 star + 4 planets + 1,000 test particle system (GPU, and the
reference's own C GR15 on one host core), the mixed-precision force error on
that system with 100,000 test particles, the 4,096-system ensemble (0.2 yr
per system) under the plain FP64 GR15, the massive-only controller and the
learned step proposer (FP64 and mixed test forces), and a 10-yr pass of the
plain GR15 (stopped at 90 s) and of the chosen form.  --configs adds the
100-yr drop-in run, the 1,000-yr mixed run with 100,000 test particles and a
1-yr ensemble pass.

The CPU baseline times the reference's own C force loop (oracle/_ref, built
from the reference sources) -- or the oracle's restatement when that is
absent -- on one host core, on a bounded sample (64 + 65,536 test
particles), and extrapolates it to the 64 + 1,048,576 workload by the N^2
loop count of the reference's loop.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

G = 4 * np.pi ** 2
FLOP_PER_INTERACTION = 20          # 3 sub, 6 for r^2, 4 for r^-3, 1 for m r^-3, 6 accumulate
FP64_PEAK_TFLOPS = 78.6            # MI355X FP64 vector peak
# the FP64 interaction issues 16 VALU ops + 1 v_rsq_f64 at quarter rate: 20
# issue slots against the 10 FMA slots the peak assumes for 20 FLOP
ISSUE_CEILING_FRAC = 10.0 / 20.0
PEAK_CLOCK_GHZ = 2.4               # the engine clock the FP64 peak is quoted at
METRIC = "interactions/s & system-yr/s at |dE/E|<=1e-10, chosen form vs plain FP64 GR15"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def interactions(n_m, n_t):
    return n_m * (n_m - 1) + n_t * n_m


def force_case(pos, m, Gc, mode, steps, warmup, each=False, roofline_ms=20.0):
    """Warm up, then time ``steps`` back-to-back launches (wall clock between
    device synchronisations, and HIP events on the launch stream).  each=True
    adds the roofline timing, independent of ``steps``: back-to-back launches
    for at least ``roofline_ms`` of device time (HIP events on the launch
    stream around the whole run), and 200 launches each bracketed by its own
    event pair (mean and median)."""
    import torch
    from abie_amd import engine
    src, sinks, n_self, _ = engine.soa_from_bodies(pos, m, Gc)
    out = tuple(torch.empty(len(m), dtype=torch.float64, device="cuda") for _ in range(3))
    for _ in range(warmup):
        engine.accel_soa(src, sinks, n_self, mode, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = engine.time_accel_soa(src, sinks, n_self, out, mode=mode, reps=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if not each:
        return wall, ms, n_self
    reps = max(200, int(np.ceil(roofline_ms / max(ms, 1e-4))))
    # the kernel's clock is read by as many clocked launches right after
    ms_run, ghz = engine.time_accel_soa(src, sinks, n_self, out, mode=mode, reps=reps, clock=True)
    _, ms_each, ms_med = engine.time_accel_soa(src, sinks, n_self, out, mode=mode, reps=200, each=True)
    # SIMD cycles per wave instruction of the interaction's instruction kinds
    # at 8 waves per SIMD: v_fma_f64, v_rsq_f64, and one rsq per 16 FMAs
    rates = {k: engine.issue_rate(k, roofline_ms) for k in ("fma", "rsq", "mix")}
    return wall, ms, n_self, {"reps": reps, "ms_back_to_back": ms_run, "ms_each_mean": ms_each,
                              "ms_each_median": ms_med, "clock_ghz": ghz, "issue_rates": rates}


def force_scan(shifts=(19, 20, 21), roofline_ms=20.0):
    """The FP64 force kernel on 64 massive + 2^k test particles: per launch
    (back-to-back, at least ``roofline_ms`` of launches), the shader clock
    read inside the kernel over as many launches, and the SIMD cycles per
    wave-source pass at that clock.  A fixed cost per launch shows as more
    cycles per pass at the smaller sizes."""
    import torch
    from abie_amd import engine, ics
    n_simd = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    mix = engine.issue_rate("mix", roofline_ms)
    out = {"issue_cycles_per_pass": 17 * mix[1], "rows": []}
    for k in shifts:
        pos, vel, m = ics.config1a(n_test=1 << k)
        src, sinks, n_self, _ = engine.soa_from_bodies(pos, m, G)
        res = tuple(torch.empty(len(m), dtype=torch.float64, device="cuda") for _ in range(3))
        ms = engine.time_accel_soa(src, sinks, n_self, res, reps=20)
        reps = max(100, int(np.ceil(roofline_ms / max(ms, 1e-4))))
        ms, ghz = engine.time_accel_soa(src, sinks, n_self, res, reps=reps, clock=True)
        passes = -(-len(m) // 256) * 4 * n_self / n_simd
        row = {"n_test": 1 << k, "us_per_launch": ms * 1e3, "launches": reps, "kernel_clock_ghz": ghz,
               "simd_cycles_per_pass": ms * 1e-3 * ghz * 1e9 / passes,
               "interactions_per_s": interactions(n_self, len(m) - n_self) / (ms * 1e-3)}
        log("force scan: %s" % row)
        out["rows"].append(row)
        del src, sinks, res
    return out


def cpu_baseline(n_test=65536, reps=3):
    """Reference C force loop on one host core (oracle/_ref), else the oracle."""
    import ctypes
    from abie_amd import ics
    pos, vel, m = ics.config1a(n_test=n_test)
    kind = "reference"
    ref = os.path.join(ROOT, "oracle", "_ref", "libabie_ref.so")
    dp = ctypes.POINTER(ctypes.c_double)
    p = np.ascontiguousarray(pos.reshape(-1))
    mm = np.ascontiguousarray(m)
    acc = np.zeros_like(p)
    r = np.zeros_like(mm)
    try:
        L = ctypes.CDLL(ref)
        f = L.ode_n_body_second_order
        f.argtypes = [dp, ctypes.c_size_t, ctypes.c_double, dp, dp, dp]
        call = lambda: f(p.ctypes.data_as(dp), len(mm), G, mm.ctypes.data_as(dp), r.ctypes.data_as(dp),
                         acc.ctypes.data_as(dp))
    except OSError:
        from oracle import oracle
        kind = "port"
        call = lambda: oracle.accel(pos, G, m)
    call()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    dt = (time.perf_counter() - t0) / reps
    n_int = interactions(64, n_test)
    # The reference's loop visits all N^2 (sink, source) pairs, massless
    # sources included (common.c:48-59), so its time per evaluation grows as
    # N^2: extrapolated to the 64 + 1,048,576 workload by that loop count.
    n_s, n_full = 64 + n_test, 64 + (1 << 20)
    dt_full = dt * (n_full / n_s) ** 2
    return {"value": interactions(64, 1 << 20) / dt_full, "unit": "interactions/s", "cores": 1, "kind": kind,
            "sample": "64 massive + %d test particles, %d evaluations, %.3f s each; value extrapolated to "
                      "64 + 1,048,576 by the N^2 loop count (%.0f s per evaluation)" % (n_test, reps, dt, dt_full),
            "sample_value": n_int / dt, "host": host_info()}


def cpu_all_cores(n_test=262144, reps=2):
    """The oracle's restatement of the same force loop with the sinks split
    over every host thread of the box (at most 16): a reported multi-core
    CPU rate beside the single-core reference baseline."""
    from abie_amd import ics
    from oracle import oracle
    k = max(1, min(16, os.cpu_count() or 1))
    pos, vel, m = ics.config1a(n_test=n_test)
    oracle.accel_threads(pos, G, m, k)
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.accel_threads(pos, G, m, k)
    dt = (time.perf_counter() - t0) / reps
    return {"value": interactions(64, n_test) / dt, "unit": "interactions/s", "cores": k, "kind": "port",
            "sample": "64 massive + %d test particles, %d evaluations, %.3f s each" % (n_test, reps, dt)}


def host_info():
    """nproc and CPU model of the host the CPU baselines run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_plummer(n=65536):
    """The reference's C force loop (oracle/_ref) on the 65,536-body Plummer
    sphere, one evaluation on one host core."""
    import ctypes
    from abie_amd import ics
    ref = os.path.join(ROOT, "oracle", "_ref", "libabie_ref.so")
    if not os.path.exists(ref):
        return None
    L = ctypes.CDLL(ref)
    dp = ctypes.POINTER(ctypes.c_double)
    f = L.ode_n_body_second_order
    f.argtypes = [dp, ctypes.c_size_t, ctypes.c_double, dp, dp, dp]
    pos, vel, m = ics.plummer(n)
    p = np.ascontiguousarray(pos.reshape(-1))
    acc = np.zeros_like(p)
    r = np.zeros(n)
    t0 = time.perf_counter()
    f(p.ctypes.data_as(dp), n, 1.0, m.ctypes.data_as(dp), r.ctypes.data_as(dp), acc.ctypes.data_as(dp))
    dt = time.perf_counter() - t0
    return {"interactions_per_s": interactions(n, 0) / dt, "wall_s": dt, "cores": 1, "kind": "reference"}


def gr15_dropin(years, force_mode="fp64", n_test=1000):
    """The 1 star + 4 planets + n_test test particle system (1,000: the
    configuration-0 system) through the reference-compatible C ABI
    (device=0), under the reference's own step controller."""
    from abie_amd import ics
    from abie_amd.clibabie import CLibABIE
    pos, vel, m = ics.config0(n_test=n_test)
    c = CLibABIE()
    c.lib.finalize_code()
    c.initialize_code(G, 0.0, len(m), 1, 1, 0.0, deviceID=0)
    c.lib.abie_hip_set_force_mode({"fp64": 0, "exact": 1, "mixed": 2}[force_mode])
    p, v = pos.reshape(-1).copy(), vel.reshape(-1).copy()
    c.set_state(p, v, m, np.zeros_like(m), len(m), G, 0.0)
    e0 = c.get_total_energy()
    t0 = time.perf_counter()
    c.integrator_gr(0.0, years, 1.0)
    wall = time.perf_counter() - t0
    t = c.get_model_time()
    c.get_state(p, v, m.copy(), np.zeros_like(m))
    e1 = c.get_total_energy()
    c.lib.abie_hip_set_force_mode(0)
    c.lib.finalize_code()
    return {"system_yr_per_s": t / wall, "model_yr": t, "wall_s": wall, "dE_E": abs((e1 - e0) / e0)}


def ensemble_sample(systems, years, force_mode, controller, proposer=None, max_seconds=120.0):
    """One pass of the batched GR15 engine over an ensemble (state resident on
    the device), bounded to ``max_seconds``; rates use the time each system
    reached."""
    from abie_amd import engine
    ens = engine.Ensemble(systems, G=G, force_mode=force_mode, controller=controller)
    if proposer is not None:
        ens.set_proposer(proposer)
    e0 = ens.energy()
    t0 = time.perf_counter()
    done = ens.run(years, max_seconds=max_seconds)
    wall = time.perf_counter() - t0
    _, _, t, st = ens.state()
    e1 = ens.energy()
    cnt = ens.counters()
    fin = ens.finish_times()
    ens.close()
    de = np.abs((e1 - e0) / e0)
    fin_ok = fin[fin >= 0]
    tail = ({"p50": float(np.median(fin_ok)), "p90": float(np.quantile(fin_ok, 0.9)),
             "p99": float(np.quantile(fin_ok, 0.99)), "max": float(fin_ok.max())} if fin_ok.size else None)
    return {"n_systems": len(systems), "model_yr_each": years, "done": bool(done), "wall_s": wall,
            "model_yr_min": float(t.min()), "model_yr_mean": float(t.mean()),
            "system_yr_per_s": float(t.sum() / wall), "evals_per_system_yr": float(cnt[:, 0].sum() / t.sum()),
            "reject_rate": float(1 - cnt[:, 2].sum() / max(cnt[:, 1].sum(), 1)),
            "dE_E_max": float(de.max()), "dE_E_p99": float(np.quantile(de, 0.99)),
            "dE_E_median": float(np.median(de)), "systems_above_1e-10": int((de > 1e-10).sum()),
            "system_yr_per_s_within_1e-10": float(t[de <= 1e-10].sum() / wall),
            "finish_s": tail}


def ensemble_report(n_sys, years, proposer):
    """Plain FP64 GR15 against the massive-only controller, the learned
    proposer, and the learned proposer with mixed-precision test forces, on
    the same ensemble."""
    from abie_amd import ics
    systems = ics.ensemble(n_sys)
    runs = [("fp64/plain", "fp64", "plain"), ("fp64/massive", "fp64", "massive")]
    if proposer is not None:
        runs += [("fp64/learned", "fp64", "learned"), ("mixed/learned", "mixed", "learned")]
    else:
        runs += [("mixed/massive", "mixed", "massive")]
    out = {}
    for name, fm, ctrl in runs:
        out[name] = ensemble_sample(systems, years, fm, ctrl, proposer if ctrl == "learned" else None)
        log("ensemble %s: %s" % (name, out[name]))
    base = out["fp64/plain"]["system_yr_per_s"]
    out["speedup_vs_plain"] = {k: v["system_yr_per_s"] / base for k, v in out.items() if k != "fp64/plain"}
    return out


def ensemble_long(n_sys, years, proposer, plain_cap=90.0, chosen_cap=240.0):
    """The plain FP64 GR15 and the chosen form (learned proposer, mixed test
    forces) over a longer horizon, where the slowest close-encounter systems
    set each form's makespan.  The plain form is stopped at ``plain_cap``
    seconds and its rate counts the model time each system reached."""
    from abie_amd import ics
    systems = ics.ensemble(n_sys)
    out = {"plain": ensemble_sample(systems, years, "fp64", "plain", max_seconds=plain_cap)}
    log("ensemble %g yr plain: %s" % (years, out["plain"]))
    if proposer is not None:
        out["chosen"] = ensemble_sample(systems, years, "mixed", "learned", proposer, max_seconds=chosen_cap)
        log("ensemble %g yr chosen (mixed/learned): %s" % (years, out["chosen"]))
        out["ratio"] = out["chosen"]["system_yr_per_s"] / out["plain"]["system_yr_per_s"]
        out["ratio_within_1e-10"] = (out["chosen"]["system_yr_per_s_within_1e-10"]
                                     / out["plain"]["system_yr_per_s_within_1e-10"])
    out["note"] = ("the configuration's 1,000-yr horizon is not measured; the plain form stops at %.0f s and "
                   "its rate counts the model time each system reached" % plain_cap)
    return out


def host_aos_case(pos, m, reps=5):
    """The reference-ABI entry point calculate_accelerations on host AoS
    arrays (PCIe copies in and out included), device 0."""
    from abie_amd.clibabie import CLibABIE
    c = CLibABIE()
    c.lib.finalize_code()
    n = len(m)
    c.initialize_code(G, 0.0, n, 1, 1, 0.0, deviceID=0)
    x = np.ascontiguousarray(pos.reshape(-1))
    v = np.zeros_like(x)
    acc = np.zeros_like(x)
    c.calculate_accelerations(x, v, n, G, m, acc=acc)  # warm: allocation, first copies
    t0 = time.perf_counter()
    for _ in range(reps):
        c.calculate_accelerations(x, v, n, G, m, acc=acc)
    dt = (time.perf_counter() - t0) / reps
    c.lib.finalize_code()
    return dt


def cpu_gr15_config0(years, chunk=None):
    """The reference's own GR15 (oracle/_ref, device -1) on the 1 star + 4
    planets + 1,000 test particle system, one host core; with ``chunk`` the
    run is split into integrator_gr calls of that many years (as the
    reference's Python driver does with store_dt), each one logged."""
    import ctypes
    from abie_amd import ics
    ref = os.path.join(ROOT, "oracle", "_ref", "libabie_ref.so")
    if not os.path.exists(ref):
        return None
    L = ctypes.CDLL(ref)
    dp = ctypes.POINTER(ctypes.c_double)
    L.initialize_code.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_size_t, ctypes.c_size_t,
                                  ctypes.c_size_t, ctypes.c_int]
    L.set_state.argtypes = [dp, dp, dp, dp, ctypes.c_size_t, ctypes.c_double, ctypes.c_double]
    L.integrator_gr.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.integrator_gr.restype = ctypes.c_int
    L.get_model_time.restype = ctypes.c_double
    pos, vel, m = ics.config0()
    p, v, mm = pos.reshape(-1).copy(), vel.reshape(-1).copy(), m.copy()
    r = np.zeros_like(mm)
    L.initialize_code(G, 0.0, len(m), 1, 1, -1)
    L.set_state(p.ctypes.data_as(dp), v.ctypes.data_as(dp), mm.ctypes.data_as(dp), r.ctypes.data_as(dp), len(m), G, 0.0)
    step = chunk or years
    t = 0.0
    t0 = time.perf_counter()
    while t < years:
        L.integrator_gr(t, min(t + step, years), 1.0)
        t = L.get_model_time()
        if chunk:
            log("cpu gr15: t = %.3f yr after %.1f s" % (t, time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    # no finalize_code: the reference initialises once per process
    # (common.c:225-227), so a later call here must find its buffers alive
    out = {"system_yr_per_s": t / wall, "model_yr": t, "wall_s": wall, "cores": 1, "kind": "reference"}
    if chunk:
        out["chunks_yr"] = chunk
    return out


def mixed_error_report(pos, m, Gc, mode="mixed"):
    """Per-test-particle relative acceleration error of a reduced-precision
    mode ("mixed" or "fp32") against FP64 (massive sinks stay FP64)."""
    import torch
    from abie_amd import engine
    src, sinks, n_self, perm = engine.soa_from_bodies(pos, m, Gc)
    a64 = torch.stack(engine.accel_soa(src, sinks, n_self, "fp64"), 1)
    amx = torch.stack(engine.accel_soa(src, sinks, n_self, mode), 1)
    tp = torch.as_tensor((m[perm] == 0), device=a64.device)
    rel = torch.linalg.norm(amx - a64, dim=1) / torch.linalg.norm(a64, dim=1)
    rel = rel[tp]
    return {"n_test": int(tp.sum()), "max_rel_err": float(rel.max()), "rms_rel_err": float(torch.sqrt((rel ** 2).mean()))}


def mixed_long_run(n_test, years, max_seconds=300.0, proposer=None):
    """GR15 of the 1 star + 4 planets system with n_test test particles under
    the massive-only controller (test particles sliced over workgroups), FP64
    vs mixed forces: |dE/E| and the test particles' final separation; with a
    proposer also the learned controller on mixed forces."""
    from abie_amd import engine, ics
    pos, vel, m = ics.config0(n_test=n_test)
    out = {}
    fin = {}
    runs = [("fp64", "fp64", "massive"), ("mixed", "mixed", "massive")]
    if proposer is not None:
        runs.append(("mixed/learned", "mixed", "learned"))
    for fm_name, fm, ctrl in runs:
        ens = engine.Ensemble([(pos, vel, m)], G=G, force_mode=fm, controller=ctrl)
        if ctrl == "learned":
            ens.set_proposer(proposer)
        e0 = ens.energy()
        t0 = time.perf_counter()
        done = ens.run(years, max_seconds=max_seconds)
        wall = time.perf_counter() - t0
        p, v, t, _ = ens.state()
        e1 = ens.energy()
        c = ens.counters()
        ens.close()
        fin[fm_name] = (p, t[0])
        out[fm_name] = {"done": bool(done), "model_yr": float(t[0]), "wall_s": wall, "dE_E": float(abs((e1[0] - e0[0]) / e0[0])),
                   "evals": int(c[0, 0]), "system_yr_per_s": float(t[0] / wall)}
    if fin["fp64"][1] == fin["mixed"][1]:
        d = np.linalg.norm(fin["mixed"][0][5:] - fin["fp64"][0][5:], axis=1) / np.linalg.norm(fin["fp64"][0][5:], axis=1)
        out["tp_position_rel_diff"] = {"max": float(d.max()), "median": float(np.median(d))}
    return out


def plummer_error_report(pos, m):
    """Per-acceleration relative difference of the fast FP64 mode from the
    exact mode (bit-identical to the reference's loop) on every sink."""
    import torch
    from abie_amd import engine
    src, sinks, n_self, _ = engine.soa_from_bodies(pos, m, 1.0)
    a = torch.stack(engine.accel_soa(src, sinks, n_self, "fp64"), 1)
    e = torch.stack(engine.accel_soa(src, sinks, n_self, "exact"), 1)
    rel = (torch.linalg.norm(a - e, dim=1) / torch.linalg.norm(e, dim=1)).cpu().numpy()
    return {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
            "p99_rel": float(np.quantile(rel, 0.99)), "frac_within_1e-14": float(np.mean(rel <= 1e-14))}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "force_c1a_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-long", action="store_true",
                    help="skip the 10-yr ensemble passes (plain capped at 90 s, chosen form)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the 4,096-system ensemble extras (for kernel traces of the GR15 engine)")
    ap.add_argument("--force-scan", action="store_true",
                    help="run only the force kernel at 64 + 2^19, 2^20, 2^21 test particles with the "
                         "kernel's own clock (cycles per wave-source pass)")
    ap.add_argument("--config2-only", action="store_true",
                    help="run only the 1 star + 4 planets + 100,000 test particle system, 100 yr, "
                         "FP64 / mixed / learned (for kernel traces)")
    ap.add_argument("--dropin-only", action="store_true",
                    help="run only the configuration-0 GR15 through the reference ABI, 10 yr "
                         "(for kernel traces of the linked-slice engine)")
    ap.add_argument("--configs", action="store_true",
                    help="also run the longer per-config measurements (config 0 full 100 yr, "
                         "config 2 1,000-yr mixed run, config 3/4 ensemble samples)")
    ap.add_argument("--proposer", default=os.path.join(ROOT, "abie_amd", "data", "proposer.npz"))
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the measurement, copy a buffer of this size 5 times "
                         "(reference traffic for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(the ensemble shards by system with no communication).", file=sys.stderr)
        sys.exit(2)

    # Everything the libraries print (the reference's C code, device printf)
    # goes to stderr; stdout carries only the JSON line.
    json_out = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)

    import torch
    from abie_amd import _lib, ics
    _lib.load()
    torch.cuda.init()

    if args.force_scan:
        json_out.write(json.dumps({"force_scan": force_scan()}) + "\n")
        json_out.flush()
        return

    if args.config2_only:
        w = np.load(args.proposer)["weights"] if os.path.exists(args.proposer) else None
        rep = mixed_long_run(100000, 100.0, proposer=w)
        log("config 2, 100 yr: %s" % rep)
        json_out.write(json.dumps({"mixed_run_config2_100yr": rep}) + "\n")
        json_out.flush()
        return

    if args.dropin_only:
        rep = gr15_dropin(10.0)
        log("gr15 drop-in: %s" % rep)
        json_out.write(json.dumps({"gr15_dropin_config0_10yr": rep}) + "\n")
        json_out.flush()
        return

    if args.ensemble_only:
        w = np.load(args.proposer)["weights"] if os.path.exists(args.proposer) else None
        rep = ensemble_report(4096, 0.2, w)
        json_out.write(json.dumps({"ensemble_config3_4096x0.2yr": rep}) + "\n")
        json_out.flush()
        return

    pos, vel, m = ics.config1a()
    n_m, n_t = int((m > 0).sum()), int((m == 0).sum())
    n_int = interactions(n_m, n_t)
    log("force workload: %d massive + %d test particles, %d interactions/step" % (n_m, n_t, n_int))
    wall, ms_event, _, rt = force_case(pos, m, G, "fp64", args.steps, args.warmup, each=True)
    value = n_int * args.steps / wall
    # roofline: algorithmic FLOP over the mean launch duration of a run of
    # back-to-back launches lasting at least 20 ms (HIP events on the launch
    # stream around the run; independent of --steps), which is what a
    # per-dispatch kernel trace averages; single launches bracketed by their
    # own events also carry the event overhead and are reported beside it
    ms_launch = rt["ms_back_to_back"]
    achieved = n_int * FLOP_PER_INTERACTION / (ms_launch * 1e-3) / 1e12
    # The 78.6 TFLOP/s peak is quoted at the 2.4 GHz peak engine clock; under a
    # full FP64 load the clock settles lower.  The kernel's own clock is read
    # inside an instance of it (per-wave shader cycles over wall time); with
    # it, the launch's SIMD cycles per wave-source pass (64 sinks against one
    # source) compare with the issue cycles of the pass's instructions (16
    # FP64 VALU + 1 v_rsq_f64, measured live as 17 instructions of the
    # one-rsq-per-16-FMA kernel).
    n_simd = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    waves = -(-len(m) // 256) * 4
    passes = waves * n_m / n_simd
    cyc_pass = ms_launch * 1e-3 * rt["clock_ghz"] * 1e9 / passes
    issue_pass = 17 * rt["issue_rates"]["mix"][1]
    ceil_sus = ISSUE_CEILING_FRAC * rt["clock_ghz"] / PEAK_CLOCK_GHZ
    log("force kernel at %.3f GHz (read inside it): %.2f SIMD cycles per wave-source pass against %.2f issue "
        "cycles (FMA %.3f, rsq %.3f cycles per wave instruction): %.3f of the issue ceiling"
        % (rt["clock_ghz"], cyc_pass, issue_pass, rt["issue_rates"]["fma"][1], rt["issue_rates"]["rsq"][1],
           issue_pass / cyc_pass))
    log("fp64: %.3f us/launch over %d back-to-back launches (%.3f over the %d timed steps), single launches "
        "%.3f us mean / %.3f us median, %.4g interactions/s"
        % (ms_launch * 1e3, rt["reps"], ms_event * 1e3, args.steps, rt["ms_each_mean"] * 1e3,
           rt["ms_each_median"] * 1e3, value))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "force kernel, 64 massive + 1,048,576 test particles (a~U[1,100] au), "
                               "one acceleration evaluation per step",
                   "n_massive": n_m, "n_test": n_t, "interactions_per_step": n_int, "mode": "fp64"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": load_traffic(),
                     "kernel": "accel_scalar_kernel<1, FP64>", "flop_per_interaction": FLOP_PER_INTERACTION,
                     "mean_launch_us": ms_launch * 1e3, "launches_timed": rt["reps"],
                     "single_launch_us_mean": rt["ms_each_mean"] * 1e3,
                     "single_launch_us_median": rt["ms_each_median"] * 1e3,
                     "issue_ceiling_frac": ISSUE_CEILING_FRAC,
                     "kernel_clock_ghz": rt["clock_ghz"],
                     "issue_ceiling_frac_at_kernel_clock": ceil_sus,
                     "simd_cycles_per_pass": cyc_pass,
                     "issue_cycles_per_pass": issue_pass,
                     "frac_of_issue_ceiling": issue_pass / cyc_pass,
                     "cycles_per_wave_instruction": {k: {"cycles": v[1], "clock_ghz": v[0]}
                                                     for k, v in rt["issue_rates"].items()},
                     "issue_ceiling_note": "16 FP64 VALU + 1 quarter-rate v_rsq_f64 = 20 four-cycle issue "
                                           "slots per wave per 64 interactions; the 78.6 TFLOP/s peak counts an "
                                           "FMA (2 FLOP) per slot, i.e. 10 slots for 20 FLOP"},
    }
    extras = {}
    if not args.no_extras:
        reps = max(10, args.steps // 20)
        pp, pv, pm = ics.plummer(65536)
        w, ms, _ = force_case(pp, pm, 1.0, "fp64", reps, 3)
        extras["plummer65536_fp64"] = {"interactions_per_s": interactions(65536, 0) * reps / w,
                                       "mean_launch_us": ms * 1e3,
                                       "frac_fp64_peak": interactions(65536, 0) * 20 / (ms * 1e-3) / 1e12
                                       / FP64_PEAK_TFLOPS,
                                       "vs_exact": plummer_error_report(pp, pm)}
        log("plummer: %s" % extras["plummer65536_fp64"])
        w, ms, _ = force_case(pos, m, G, "mixed", args.steps, args.warmup)
        # the mixed mode's test-particle terms are FP64 instructions (uncorrected
        # rsqrt seed), so its work is priced against the FP64 vector peak
        extras["config1a_mixed"] = {"interactions_per_s": n_int * args.steps / w, "mean_launch_us": ms * 1e3,
                                    "frac_fp64_peak": n_int * 20 / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        # the FP32 formulation of the test-particle terms, measured beside it
        w, ms, _ = force_case(pos, m, G, "fp32", args.steps, args.warmup)
        extras["config1a_fp32"] = {"interactions_per_s": n_int * args.steps / w, "mean_launch_us": ms * 1e3,
                                   "frac_fp64_peak": n_int * 20 / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        w, ms, _ = force_case(pos, m, G, "exact", reps, 3)
        extras["config1a_exact"] = {"interactions_per_s": n_int * reps / w, "mean_launch_us": ms * 1e3}
        log("mixed/fp32/exact: %s %s %s" % (extras["config1a_mixed"], extras["config1a_fp32"], extras["config1a_exact"]))
        dt = host_aos_case(pos, m)
        extras["config1a_host_aos_pcie"] = {"interactions_per_s": n_int / dt, "ms_per_call": dt * 1e3,
                                            "note": "calculate_accelerations on host AoS arrays: pageable "
                                                    "copies over PCIe in and out, reorder kernels, force kernel"}
        log("host AoS: %s" % extras["config1a_host_aos_pcie"])
        extras["gr15_dropin_config0_10yr"] = gr15_dropin(10.0)
        log("gr15 drop-in: %s" % extras["gr15_dropin_config0_10yr"])
        p2, _, m2 = ics.config0(n_test=100000)
        extras["mixed_error_config2"] = mixed_error_report(p2, m2, G)
        extras["mixed_error_config1a"] = mixed_error_report(pos, m, G)
        extras["fp32_error_config2"] = mixed_error_report(p2, m2, G, "fp32")
        extras["fp32_error_config1a"] = mixed_error_report(pos, m, G, "fp32")
        log("mixed error: C2 %s, C1a %s; fp32 error: C2 %s, C1a %s" % (extras["mixed_error_config2"], extras["mixed_error_config1a"], extras["fp32_error_config2"], extras["fp32_error_config1a"]))
        extras["gr15_config0_cpu_reference_1yr"] = cpu_gr15_config0(1.0) if not args.no_cpu else None
        log("gr15 cpu reference: %s" % extras["gr15_config0_cpu_reference_1yr"])
        w = np.load(args.proposer)["weights"] if os.path.exists(args.proposer) else None
        extras["ensemble_config3_4096x0.2yr"] = ensemble_report(4096, 0.2, w)
        if not args.no_long:
            extras["ensemble_config3_4096x10yr"] = ensemble_long(4096, 10.0, w)
    if args.configs:
        extras["gr15_dropin_config0_100yr"] = gr15_dropin(100.0)
        log("config 0, 100 yr: %s" % extras["gr15_dropin_config0_100yr"])
        # 100,005 bodies through the reference ABI: linked slices under the
        # reference's controller (test particles in its error norm)
        extras["gr15_dropin_config2_plain_2yr"] = gr15_dropin(2.0, n_test=100000)
        log("config 2 through the reference ABI, 2 yr: %s" % extras["gr15_dropin_config2_plain_2yr"])
        w = np.load(args.proposer)["weights"] if os.path.exists(args.proposer) else None
        extras["mixed_run_config2"] = mixed_long_run(100000, 1000.0, proposer=w)
        log("config 2 run: %s" % extras["mixed_run_config2"])
        extras["ensemble_config3_4096x1yr"] = ensemble_report(4096, 1.0, w)
        if not args.no_cpu:
            extras["cpu_reference_plummer65536"] = cpu_plummer()
            log("cpu plummer: %s" % extras["cpu_reference_plummer65536"])
            extras["cpu_reference_gr15_config0_100yr"] = cpu_gr15_config0(100.0, chunk=10.0)
            log("cpu gr15 100 yr: %s" % extras["cpu_reference_gr15_config0_100yr"])
    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        extras["copy_calibration_bytes"] = nel * 8
    out["extras"] = extras
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        log("cpu: %s" % out["cpu_baseline"])
        out["extras"]["cpu_restatement_all_cores"] = cpu_all_cores()
        log("cpu, all cores: %s" % out["extras"]["cpu_restatement_all_cores"])
    else:
        out["cpu_baseline"] = None
    import ctypes
    sys.stdout.flush()
    ctypes.CDLL(None).fflush(None)
    json_out.write(json.dumps(out) + "\n")
    json_out.flush()


if __name__ == "__main__":
    main()
