"""Benchmark of the GRAPE-6 force path on one MI355X.

Workload: N = 2^20 Plummer sphere, Salpeter masses on [0.1, 100] normalised
to M = 1, eps2 = 1e-8 (the large single-system configuration).  A step is
one full-N force evaluation (acc, jerk, pot and nearest neighbour of every
particle against all 2^20 j-slots, 2^40 pair interactions) on data already
resident in HBM, through the library's device-resident entry point
g6x_calc_self.  Both paths are timed: the plain FP64 path and the mixed
double-single/FP32 path; `value` is the mixed path, the FP64 path and the
ratio are reported beside it.

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu]

Prints one JSON line.  Kernel time comes from HIP events the library
records on its own stream around every force launch.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 60.0          # GRAPE-6 convention (src/amuse_ph4/src/jdata.cc:1038)
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}   # MI355X vector FP64 / FP32
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_system(n: int, seed: int = 20240):
    from amuse_amd import ic
    rng = np.random.Generator(np.random.PCG64(seed))
    _, pos, vel = ic.plummer_raw(n, rng)
    pos, vel = ic.to_nbody_units(pos, vel)
    m = ic.salpeter(n, rng, 0.1, 100.0, -2.35)
    m = m / m.sum()
    pos, vel = ic.center(m, pos, vel)
    return m, pos, vel


def upload(g6, m, pos, vel):
    n = m.size
    z = np.zeros((n, 3))
    g6.set_j_particles(np.arange(n, dtype=np.int32), np.arange(1, n + 1, dtype=np.int32),
                       np.zeros(n), np.zeros(n), m, z, z, z, vel, pos)
    g6.set_ti(0.0)


def time_path(g6, prec: int, n: int, eps2: float, steps: int, warmup: int):
    g6.set_precision(prec)
    for _ in range(warmup):
        g6.calc_self(n, 0, n, eps2, nn=True)
    g6.sync()
    g6.timing(True)
    t0 = time.perf_counter()
    for s in range(steps):
        g6.calc_self(n, 0, n, eps2, nn=True)
        g6.sync()
        log(f"  prec={prec} step {s + 1}/{steps} done")
    wall = time.perf_counter() - t0
    tm = g6.get_timing()
    g6.timing(False)
    res = g6.fetch(n)
    return wall, tm, res


def cpu_baseline(m, pos, vel, eps2: float, threads: int, per_thread: int):
    import oracle
    n = m.size
    ni = threads * per_thread
    z = np.zeros((n, 3))
    ids = np.arange(1, n + 1, dtype=np.int32)
    sec, acc, jerk, pot, nn = oracle.baseline_run(threads, per_thread, n, ni, 0.0, eps2, ids,
                                                  np.zeros(n), m, pos, vel, z, z, z)
    return sec, ni, (acc, jerk, pot, nn)


def _read(path: str) -> str | None:
    try:
        with open(path) as f:
            return f.read()
    except OSError:
        return None


def host_info() -> dict:
    """The host the CPU baselines run on: logical CPUs, the ones this process
    may run on (affinity), the cgroup CPU quota, CPU model and clock."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    quota = None
    cg = _read("/sys/fs/cgroup/cpu.max")                   # cgroup v2: "quota period"
    if cg and cg.split()[0] != "max":
        quota = float(cg.split()[0]) / float(cg.split()[1])
    else:
        q = _read("/sys/fs/cgroup/cpu/cpu.cfs_quota_us")   # cgroup v1
        p = _read("/sys/fs/cgroup/cpu/cpu.cfs_period_us")
        if q and p and int(q) > 0:
            quota = int(q) / int(p)
    info["cgroup_cpu_quota"] = quota
    cpuinfo = _read("/proc/cpuinfo") or ""
    models = [ln.split(":", 1)[1].strip() for ln in cpuinfo.splitlines()
              if ln.startswith("model name")]
    mhz = [float(ln.split(":", 1)[1]) for ln in cpuinfo.splitlines() if ln.startswith("cpu MHz")]
    info["cpu_model"] = models[0] if models else None
    info["cpu_mhz_now_mean"] = round(sum(mhz) / len(mhz), 1) if mhz else None
    fmax = _read("/sys/devices/system/cpu/cpu0/cpufreq/cpuinfo_max_freq")
    info["cpu_mhz_max"] = int(fmax) / 1000 if fmax else None
    return info


def usable_cores(info: dict) -> int:
    """Threads the CPU baselines use: every CPU of the affinity mask, capped
    by the cgroup quota when one is set (threads beyond it are throttled)."""
    n = info["affinity"]
    if info["cgroup_cpu_quota"]:
        n = min(n, max(1, int(info["cgroup_cpu_quota"])))
    return n


def load_traffic(kernel: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(kernel)


def copy_calibration() -> int:
    """A plain device-to-device copy of exactly 1 GiB (torch's copy kernel),
    for scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
    return 0


def blocks_study(args) -> int:
    """Force-only study: N = 65536 Plummer (equal masses, eps2 = 1e-4), every
    j with its own time t_j = t_i - U(0,1) 2^-U{3..10}, acc and jerk from one
    FP64 pass, i-sets = the first 2^k particles predicted to t_i, k = 0..16.
    Per k: kernel time of both paths, wall time of one ABI call (firsthalf +
    lasthalf2 from host arrays, i.e. PCIe and synchronisation included), and
    per-particle errors against the 80-bit oracle on (up to) 256 sampled i."""
    import oracle
    from amuse_amd import ic
    from amuse_amd.g6 import G6
    n, eps2 = 65536, 1e-4
    rng = np.random.Generator(np.random.PCG64(65536))
    m, x, v = ic.plummer(n, 1)
    ids = np.arange(1, n + 1, dtype=np.int32)
    g6 = G6(0)
    g6.open()
    z = np.zeros((n, 3))
    g6.set_j_particles(np.arange(n), ids, np.zeros(n), 0.0, m, z, z, z, v, x)
    g6.set_ti(0.0)
    g6.calc_self(n, 0, n, eps2)
    acc, jerk, _, _ = g6.fetch(n)
    ti = 1.0
    tj = ti - rng.uniform(0, 1, n) * 2.0 ** -rng.integers(3, 11, n).astype(float)
    a2, j6 = acc / 2, jerk / 6
    g6.set_j_particles(np.arange(n), ids, tj, 0.0, m, z, j6, a2, v, x)
    g6.set_ti(ti)
    xp, vp = oracle.predict(ti, tj, x, v, a2, j6)
    rows = []
    for k in range(0, 17):
        ni = 1 << k
        sel = np.arange(ni)
        samp = sel if ni <= 256 else np.linspace(0, ni - 1, 256).astype(int)
        ex = oracle.force_ld(ids[samp], xp[samp], vp[samp], eps2, ids, m, xp, vp)
        row = {"k": k, "ni": ni}
        for prec, name in ((0, "fp64"), (1, "mixed")):
            g6.set_precision(prec)
            g6.calc(n, ids[sel], xp[sel], vp[sel], eps2)     # warm-up
            g6.timing(True)
            reps = max(1, min(20, (1 << 22) // ni))
            # no collector pass inside the timed calls (one of ~45 ms once
            # landed in the 20 calls of one block size)
            gc.collect()
            gc.disable()
            w0 = time.perf_counter()
            for _ in range(reps):
                out = g6.calc(n, ids[sel], xp[sel], vp[sel], eps2)
            call_us = (time.perf_counter() - w0) / reps * 1e6
            gc.enable()
            tm = g6.get_timing()
            g6.timing(False)
            kms = tm["force_ms"] / tm["force_launches"]
            pms = tm["predict_ms"]
            errs = {}
            for val, e, sm, q in ((out[0][samp], ex[0], ex[3], "acc"),
                                  (out[1][samp], ex[1], ex[4], "jerk"),
                                  (out[2][samp], ex[2], ex[5], "pot")):
                d = np.abs(val - e).reshape(len(samp), -1)
                d = np.sqrt((d * d).sum(1))
                mag = np.abs(e).reshape(len(samp), -1)
                mag = np.sqrt((mag * mag).sum(1))
                rel, cond = d / mag, d / sm
                errs[q] = {"rel_max": float(rel.max()), "rel_p99": float(np.quantile(rel, .99)),
                           "cond_max": float(cond.max())}
            row[name] = {"kernel_ms": kms, "gpairs_per_s": ni * n / kms / 1e6,
                         "reduce_us": tm["reduce_ms"] / tm["force_launches"] * 1e3,
                         "abi_call_us": call_us,
                         "predict_ms_total": pms, "errors": errs}
        rows.append(row)
        log(f"  k={k}: fp64 {row['fp64']['gpairs_per_s']:.1f} mixed "
            f"{row['mixed']['gpairs_per_s']:.1f} Gpairs/s; acc rel max fp64 "
            f"{row['fp64']['errors']['acc']['rel_max']:.2e} mixed "
            f"{row['mixed']['errors']['acc']['rel_max']:.2e}")
    g6.close()
    cpu = None
    if not args.no_cpu:
        # the g6lib.c restatement on the host cores, same j-set and i-blocks:
        # each thread's i-share in one call, and (up to 1024 i) one i per call
        # as the reference's g6_npipes() = 1 (lib/g6lib/g6lib.c:228-231),
        # every call predicting the whole j-set as g6lib does
        host = host_info()
        threads = args.cpu_threads or usable_cores(host)
        cpu = {"kind": "port", "host": host, "rows": []}
        for row in rows:
            ni = row["ni"]
            th = min(threads, ni)
            crow = {"k": row["k"], "ni": ni, "threads": th}
            for label, npipes in (("npipes_share", max(1, ni // th)), ("npipes_1", 1)):
                if label == "npipes_1" and ni > 1024:
                    continue
                sec = oracle.baseline_run(th, npipes, n, ni, ti, eps2, ids, tj, m, x, v, a2, j6)[0]
                crow[label] = {"seconds": sec, "gpairs_per_s": ni * n / sec / 1e9,
                               "speedup_of_gpu_abi_call": {
                                   p: sec / (row[p]["abi_call_us"] * 1e-6) for p in ("fp64", "mixed")}}
            cpu["rows"].append(crow)
            log(f"  cpu k={row['k']}: {crow['npipes_share']['gpairs_per_s']:.3f} Gpairs/s on {th} threads")
    print(json.dumps({"workload": "i-blocks 2^k (k=0..16) vs N=65536 Plummer, per-j times, "
                                  "eps2=1e-4", "rows": rows, "cpu_baseline": cpu}))
    return 0


def tile_sweep(args) -> int:
    """Force-kernel time per tile shape for ni = 65536 against nj = N."""
    from amuse_amd.g6 import G6
    n, ni = args.n, min(65536, args.n)
    m, pos, vel = make_system(n)
    g6 = G6(0)
    g6.open()
    upload(g6, m, pos, vel)
    rows = []
    for prec in (0, 1):
        g6.set_precision(prec)
        for njs, ipl, smem in ((1, 2, 1), (1, 1, 1), (1, 2, 0), (1, 1, 0), (2, 2, 0), (4, 1, 0)):
            g6.set_tiles(njs, ipl, 0, smem)
            g6.calc_self(n, 0, ni, args.eps2)
            g6.sync()
            g6.timing(True)
            for _ in range(3):
                g6.calc_self(n, 0, ni, args.eps2)
            tm = g6.get_timing()
            g6.timing(False)
            k = tm["force_ms"] / tm["force_launches"]
            rows.append({"prec": prec, "njs": njs, "ipl": ipl, "smem": smem, "kernel_ms": k,
                         "gpairs_per_s": ni * n / k / 1e6})
            log(f"  {rows[-1]}")
    g6.set_tiles(0, 0)
    g6.close()
    print(json.dumps({"workload": f"tile sweep ni={ni} nj={n}", "rows": rows}))
    return 0


def evolve_blocks(m, pos, vel, eps2, eta, nblocks, precision, device=False):
    """ph4-style run, host-side driver through the GRAPE-6 ABI or the
    device-resident driver: commit (full-N initial force, FP64) then nblocks
    block steps with forces in `precision`; energies from FP64 passes.
    Returns (dE/E, pairs of the block steps, seconds of the block steps,
    block and particle step counts)."""
    from amuse_amd.blocks import DeviceHermite4
    from amuse_amd.hermite import Hermite4
    # the initial forces and steps in FP64 for both paths, so that both
    # start from the same block structure
    h = DeviceHermite4(precision=0, steps_per_graph=8) if device else Hermite4(precision=0)
    h.commit(m, pos, vel, eps2=eps2, eta=eta)
    e0 = sum(h.energy())
    h.set_precision(precision)
    if device:
        # capture the step graph outside the timed region: evolving to the
        # current time takes no step
        h.evolve(h.time)
    c0 = h.counters()
    t0 = time.perf_counter()
    h.steps(nblocks)
    dt = time.perf_counter() - t0
    c1 = h.counters()
    h.set_precision(0)
    e1 = sum(h.energy())
    h.close()
    ps = c1["particle_steps"] - c0["particle_steps"]
    return {"dE_over_E": (e1 - e0) / abs(e0), "E0": e0, "E1": e1,
            "pairs": float(ps) * m.size, "seconds": dt,
            "gpairs_per_s": float(ps) * m.size / dt / 1e9,
            "block_steps": c1["block_steps"] - c0["block_steps"], "particle_steps": ps}


def ph4_c0(args) -> int:
    """Configuration c0: Plummer N = 1024, eps2 = 1e-4, eta = 0.01 to t = 10
    with the block-step driver on libg6.so in FP64 (host-side driver over the
    GRAPE-6 ABI, and the device-resident driver), energy error every 0.5
    (FP64 passes) beside the reference ph4 integrator's run on the CPU oracle
    (tests/golden/ph4_c0_energy.npz).  Mixed forces: at eta = 0.01 with a
    block budget of 4x the FP64 run's (the step rule's third force difference
    drops below FP32 rounding there), and FP64 and mixed at ph4's default
    eta = 0.14 to t = 10 (host-side and device-resident drivers)."""
    from amuse_amd import ic
    from amuse_amd.hermite import Hermite4
    gold = np.load(os.path.join(ROOT, "tests", "golden", "ph4_c0_energy.npz"))
    m, x, v = ic.plummer(int(gold["n"]), int(gold["seed"]))
    out = {"workload": "c0: Plummer N=1024, eps2=1e-4, eta=0.01, t=0..10",
           "t": gold["t"].tolist(),
           "reference_ph4_on_cpu_oracle": {
               "dE_over_E": ((gold["E"] - gold["E"][0]) / abs(gold["E"][0])).tolist(),
               "block_steps": gold["blocks"].tolist(), "particle_steps": gold["steps"].tolist()}}

    def one(prec, eta, budget_per_out=None, device=False):
        if device:
            from amuse_amd.blocks import DeviceHermite4
            h = DeviceHermite4(precision=prec)
        else:
            h = Hermite4(precision=prec)
        h.commit(m, x, v, eps2=float(gold["eps2"]), eta=eta)
        E, B, S, T = [], [], [], []
        t0 = time.perf_counter()
        reached = True
        for i, t in enumerate(gold["t"]):
            h.set_precision(prec)
            budget = 0 if budget_per_out is None else int(budget_per_out[i]) + 1
            reached = h.evolve(float(t), budget)
            h.set_precision(0)
            E.append(sum(h.energy()))
            c = h.counters()
            B.append(c["block_steps"]); S.append(c["particle_steps"]); T.append(h.time)
            log(f"  {'device' if device else 'host'} driver prec={prec} eta={eta} t={t:.2f} "
                f"dE/E={(E[-1] - E[0]) / abs(E[0]):.3e} blocks={B[-1]} reached={reached}")
            if not reached:
                break
        wall = time.perf_counter() - t0
        h.close()
        E = np.array(E)
        return {"driver": "device" if device else "host", "eta": eta,
                "dE_over_E": ((E - E[0]) / abs(E[0])).tolist(), "block_steps": B,
                "particle_steps": S, "reached_all": reached, "time_reached": T[-1],
                "wall_s": wall, "gpairs_per_s": float(S[-1]) * m.size / wall / 1e9}

    eta0 = float(gold["eta"])
    out["fp64"] = one(0, eta0)
    out["fp64_device_driver"] = one(0, eta0, device=True)
    budget = 4 * np.diff(np.concatenate([[0], gold["blocks"]]))
    out["mixed"] = one(1, eta0, budget)
    out["fp64_eta0.14"] = one(0, 0.14)
    out["mixed_eta0.14"] = one(1, 0.14)
    out["fp64_eta0.14_device_driver"] = one(0, 0.14, device=True)
    out["mixed_eta0.14_device_driver"] = one(1, 0.14, device=True)
    print(json.dumps(out))
    return 0


def driver_study(args) -> int:
    """Device-resident block-step driver: Plummer N in --driver-n (default
    1024, 16384, 65536), eps2 = 1e-4, eta = 0.01, FP64; time per block step
    from t = 1/64 to --driver-t (default 1/16)."""
    from amuse_amd import ic
    from amuse_amd.blocks import DeviceHermite4
    rows = []
    for n in [int(q) for q in args.driver_n.split(",")]:
        m, x, v = ic.plummer(n, 1024)
        d = DeviceHermite4(precision=0)
        d.commit(m, x, v, eps2=1e-4, eta=0.01)
        d.evolve(1 / 64)
        c0 = d.counters()
        t0 = time.perf_counter()
        d.evolve(args.driver_t)
        dt = time.perf_counter() - t0
        c1 = d.counters()
        d.close()
        nb = c1["block_steps"] - c0["block_steps"]
        ps = c1["particle_steps"] - c0["particle_steps"]
        rows.append({"n": n, "block_steps": nb, "us_per_block": dt / nb * 1e6,
                     "mean_block": ps / nb, "gpairs_per_s": ps * n / dt / 1e9})
        log(f"  driver n={n}: {rows[-1]}")
    print(json.dumps({"workload": "device-resident block steps, Plummer, eps2=1e-4, eta=0.01, "
                                  "FP64", "rows": rows}))
    return 0


def c4_evolve(args) -> int:
    """Configuration c4 integrated: N = 2^20 Plummer/Salpeter (eps2 = 1e-8),
    ph4's block steps with the device-resident driver from t = 0 to
    --c4-t (default 1/64) at --eta (ph4's default 0.14), once with FP64
    forces and once with mixed ones (initial forces and steps in FP64 for
    both); energies from FP64 passes.  Reports |dE/E| of each path and their
    ratio, block and particle steps, wall time and pair interactions/s."""
    n, eps2 = args.n, args.eps2
    m, pos, vel = make_system(n)
    print(json.dumps(c4_evolve_run(m, pos, vel, eps2, args.eta, args.c4_t)))
    return 0


def c4_evolve_run(m, pos, vel, eps2, eta, t_end, nout=16):
    from amuse_amd.blocks import DeviceHermite4
    n = m.size
    res = {"workload": f"c4 evolve: N={n} Plummer/Salpeter, eps2={eps2}, eta={eta}, "
                       f"t=0..{t_end}, device-resident driver"}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        d = DeviceHermite4(precision=0)
        d.commit(m, pos, vel, eps2=eps2, eta=eta)
        e0 = sum(d.energy())
        d.set_precision(prec)
        c0 = d.counters()
        t0 = time.perf_counter()
        for k in range(1, nout + 1):
            d.evolve(t_end * k / nout)
            c = d.counters()
            log(f"  {name}: t={c['system_time']:.6f} blocks={c['block_steps']} "
                f"{time.perf_counter() - t0:.1f} s")
        wall = time.perf_counter() - t0
        c1 = d.counters()
        d.set_precision(0)
        e1 = sum(d.energy())
        d.close()
        ps = c1["particle_steps"] - c0["particle_steps"]
        res[name] = {"dE_over_E": (e1 - e0) / abs(e0), "E0": e0, "E1": e1,
                     "block_steps": c1["block_steps"] - c0["block_steps"],
                     "particle_steps": ps, "wall_s": wall,
                     "gpairs_per_s": float(ps) * n / wall / 1e9}
        log(f"  {name}: {res[name]}")
    res["energy_error_ratio_mixed_over_fp64"] = abs(res["mixed"]["dE_over_E"]) / max(
        abs(res["fp64"]["dE_over_E"]), 1e-300)
    res["speed_ratio_mixed_over_fp64"] = res["fp64"]["wall_s"] / res["mixed"]["wall_s"]
    return res


def worker_study(args) -> int:
    """The device-resident integrator out of process through AMUSE's socket
    protocol (amuse_amd/lib/g6_worker): round trip of a scalar call, get_state
    of N = 65536 particles in one vectorised call, and ph4's c0 loop (Plummer
    N = 1024, eps2 = 1e-4, eta = 0.01, t = 0..1/8 in 8 evolve_model calls)
    through the worker against the same run in process."""
    from amuse_amd import ic
    from amuse_amd.blocks import DeviceHermite4
    from amuse_amd.worker import Ph4Worker
    res = {"workload": "socket worker (AMUSE message protocol) around the device driver"}
    w = Ph4Worker()
    try:
        w.call("initialize_code")
        gc.collect()
        gc.disable()    # no collector pass inside the 2000 round trips
        t0 = time.perf_counter()
        for _ in range(2000):
            w.call("get_eps2")
        res["scalar_call_us"] = (time.perf_counter() - t0) / 2000 * 1e6
        gc.enable()
        log(f"  worker: scalar call {res['scalar_call_us']:.1f} us")
        n = 65536
        rng = np.random.Generator(np.random.PCG64(9))
        m, x, v = np.full(n, 1.0 / n), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        r, out = w.call("new_particle", m, *x.T, *v.T, np.zeros(n), -1)
        ids = out["index_of_the_particle"]
        t0 = time.perf_counter()
        for _ in range(10):
            w.call("get_state", ids)
        dt = (time.perf_counter() - t0) / 10
        nbytes = n * (4 + 4 + 8 * 8)  # index in, result + 8 doubles out
        res["get_state_65536"] = {"ms": dt * 1e3, "MB_per_s": nbytes / dt / 1e6}
        log(f"  worker: get_state of {n}: {res['get_state_65536']}")
        w.call("cleanup_code")
        m, x, v = ic.plummer(1024, 1024)
        w.call("set_eps2", 1e-4)
        w.call("set_eta", 0.01)
        w.call("new_particle", m, *x.T, *v.T, np.zeros(1024), -1)
        t0 = time.perf_counter()
        if w.call("commit_particles")[0] != 0:
            raise RuntimeError("worker: commit_particles failed (no GPU?)")
        for k in range(1, 9):
            if w.call("evolve_model", k / 64)[0] != 0:
                raise RuntimeError("worker: evolve_model failed")
        e_w = w.call("get_kinetic_energy")[1]["kinetic_energy"] + \
            w.call("get_potential_energy")[1]["potential_energy"]
        t_w = time.perf_counter() - t0
        blocks = w.call("get_block_steps")[1]["block_steps"]
    finally:
        w.stop()
    d = DeviceHermite4(precision=0)
    t0 = time.perf_counter()
    d.commit(m, x, v, eps2=1e-4, eta=0.01)
    for k in range(1, 9):
        d.evolve(k / 64)
    e_d = sum(d.energy())
    t_d = time.perf_counter() - t0
    d.close()
    res["c0_t_0_to_1_8"] = {"block_steps": blocks, "worker_s": t_w, "in_process_s": t_d,
                            "energy_equal": bool(e_w == e_d)}
    log(f"  worker: c0 to t=1/8: {res['c0_t_0_to_1_8']}")
    print(json.dumps(res))
    return 0


def clusters_c2(args) -> int:
    """Configuration c2: 4096 clusters, N log-uniform in [16, 2048], Kroupa
    0.08-20 normalised to M = 1 per cluster, eps2 = 1e-6, all to t = 1 in
    mixed precision with autotuned tiles (lanes per i and j per tile), the
    step kernels launched one by one (--c2-mode eager, the default),
    as hipGraph-captured steps (--c2-mode graph) or the whole run in one
    persistent cooperative launch (--c2-mode persistent); the same ensemble
    in FP64 for the energy comparison."""
    from amuse_amd.ensemble import Ensemble, make_ensemble
    nc = 4096 if args.clusters <= 0 else args.clusters
    counts, m, x, v = make_ensemble(nc, 2048)
    res = {"workload": f"c2: {nc} clusters, N log-uniform 16..2048, Kroupa, eps2=1e-6, t=1",
           "particles": int(counts.sum()), "clusters": nc}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        ens = Ensemble(counts, m, x, v, eps2=1e-6, dt_param=args.dt_param_c2, precision=prec)
        tuned = ens.autotune()
        tiles = ens.tiles()
        jtiles = ens.jtiles()
        e0 = ens.energy()
        ens.start(1.0)
        limit = args.max_rounds if args.max_rounds > 0 else 4000000
        t0 = time.perf_counter()
        rounds = 0
        if args.c2_mode == "persistent":
            log(f"  {name}: one persistent launch")
            rounds = ens.run(persistent=True, max_rounds=limit)
        else:
            chunk = 4096  # step rounds per call, so progress is printed
            while rounds < limit:
                r = ens.run(steps_per_graph=16, use_graph=args.c2_mode == "graph",
                            max_rounds=min(chunk, limit - rounds))
                rounds += r
                log(f"  {name}: {rounds} step rounds, {time.perf_counter() - t0:.1f} s")
                if r < chunk:
                    break
        wall = time.perf_counter() - t0
        e1 = ens.energy()
        _, _, tc, ns = ens.state()
        ens.close()
        pairs = float((ns.astype(np.float64) * counts.astype(np.float64) ** 2).sum())
        de = np.abs((e1 - e0) / e0)
        res[name] = {"wall_s": wall, "step_rounds": int(rounds),
                     "cluster_steps_total": int(ns.sum()), "finished": int((tc >= 1.0).sum()),
                     "gpairs_per_s": pairs / wall / 1e9, "tiles_njs_per_size_class": tiles,
                     "tiles_j_per_size_class": jtiles, "mode": args.c2_mode,
                     "autotune_timings": int(tuned),
                     "dE_over_E_median": float(np.median(de)),
                     "dE_over_E_p90": float(np.quantile(de, 0.9)),
                     "dE_over_E_max": float(de.max())}
        log(f"  {name}: {res[name]}")
    res["mixed_over_fp64_speed"] = res["mixed"]["gpairs_per_s"] / res["fp64"]["gpairs_per_s"]
    print(json.dumps(res))
    return 0


def adjoint_c3(args) -> int:
    """Configuration c3: N = 16384 Plummer, equal masses, eps2 = 1e-4,
    256 fixed Hermite steps of 1/256, checkpoint every 16; loss on the final
    10/50/90 % Lagrangian radii (targets 0.9x initial); gradients w.r.t. x0,
    v0, m in FP32, checked against the FP64 pipeline and central finite
    differences."""
    import torch
    from amuse_amd import ic
    from amuse_amd.diff import hermite_integrate, lagrangian_radii, radii_loss
    n = args.n_c3
    m, x, v = ic.plummer(n, 3)
    dev = torch.device("cuda")
    nsteps, dt, eps2, ck = args.steps_c3, 1.0 / 256, 1e-4, 16

    def run(dtype, grads=True):
        xx = torch.tensor(x, dtype=dtype, device=dev, requires_grad=grads)
        vv = torch.tensor(v, dtype=dtype, device=dev, requires_grad=grads)
        mm = torch.tensor(m, dtype=dtype, device=dev, requires_grad=grads)
        with torch.no_grad():
            targets = 0.9 * lagrangian_radii(xx, mm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xf, _ = hermite_integrate(xx, vv, mm, dt, nsteps, eps2, checkpoint_every=ck)
        loss = radii_loss(xf, mm, targets)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if grads:
            loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = (xx.grad, vv.grad, mm.grad) if grads else None
        return float(loss.detach()), g, t1 - t0, t2 - t1

    run(torch.float32)   # warm-up
    L32, g32, tf32, tb32 = run(torch.float32)
    log(f"  c3 fp32: loss {L32:.6e} forward {tf32:.2f} s backward {tb32:.2f} s")
    L64, g64, tf64, tb64 = run(torch.float64)
    log(f"  c3 fp64: loss {L64:.6e} forward {tf64:.2f} s backward {tb64:.2f} s")
    pairs_fwd = float(n) * n * (nsteps + 1)
    # the backward recomputes each checkpointed segment (forward pairs again)
    # and runs one adjoint pass per force evaluation
    rel = [float(torch.linalg.vector_norm(a.double() - b) / torch.linalg.vector_norm(b))
           for a, b in zip(g32, g64)]
    # central finite differences of the FP64 loss along two random directions
    gen = torch.Generator(device=dev).manual_seed(11)
    fd = []
    for which in (0, 1):
        d = torch.randn((n, 3), dtype=torch.float64, device=dev, generator=gen)
        h = 1e-6
        base = [torch.tensor(q, dtype=torch.float64, device=dev) for q in (x, v, m)]

        def loss_at(sign):
            b = [t.clone() for t in base]
            b[which] = b[which] + sign * h * d
            with torch.no_grad():
                tg = 0.9 * lagrangian_radii(torch.tensor(x, dtype=torch.float64, device=dev),
                                            torch.tensor(m, dtype=torch.float64, device=dev))
                xf, _ = hermite_integrate(b[0], b[1], b[2], dt, nsteps, eps2, checkpoint_every=0)
                return float(radii_loss(xf, b[2], tg))
        num = (loss_at(1) - loss_at(-1)) / (2 * h)
        an = float((g64[which] * d).sum())
        fd.append({"wrt": ["x0", "v0"][which], "analytic": an, "finite_diff": num,
                   "rel_err": abs(an - num) / max(abs(num), 1e-300)})
        log(f"  c3 finite differences: {fd[-1]}")
    out = {"workload": f"c3: Plummer N={n}, {nsteps} Hermite steps dt=1/256, checkpoint {ck}, "
                       "Lagrangian-radius loss",
           "loss_fp32": L32, "loss_fp64": L64,
           "fp32": {"forward_s": tf32, "backward_s": tb32,
                    "forward_gpairs_per_s": pairs_fwd / tf32 / 1e9},
           "fp64": {"forward_s": tf64, "backward_s": tb64,
                    "forward_gpairs_per_s": pairs_fwd / tf64 / 1e9},
           "grad_rel_err_fp32_vs_fp64": {"x0": rel[0], "v0": rel[1], "m": rel[2]},
           "finite_difference_check_fp64": fd}
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--eps2", type=float, default=1e-8)
    ap.add_argument("--eta", type=float, default=0.14,
                    help="c4 block steps: Aarseth eta (ph4's default timestep_parameter, "
                         "src/amuse_ph4/interface.py:514-520; at 0.01 FP32 force rounding "
                         "exceeds the step rule's third force difference)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-per-thread", type=int, default=256)
    ap.add_argument("--clusters", type=int, default=0, help="c2: number of clusters")
    ap.add_argument("--c2-mode", choices=["persistent", "graph", "eager"], default="eager",
                    help="c2: the step kernels launched one by one (fastest measured: they run "
                         "back to back), graph-replayed steps, or one persistent launch for the "
                         "run (grid barriers between phases)")
    ap.add_argument("--dt-param-c2", type=float, default=0.03,
                    help="c2: hermite dt_param (step = dt_param x collision time)")
    ap.add_argument("--max-rounds", type=int, default=0)
    ap.add_argument("--n-c3", type=int, default=16384)
    ap.add_argument("--driver-n", default="1024,16384,65536", help="driver: system sizes")
    ap.add_argument("--driver-t", type=float, default=1 / 16, help="driver: end time")
    ap.add_argument("--steps-c3", type=int, default=256)
    ap.add_argument("--c4-t", type=float, default=1 / 64, help="c4evolve: end time")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--workload", default="full",
                    choices=["full", "blocks", "tiles", "ph4", "clusters", "adjoint", "driver",
                             "worker", "c4evolve"],
                    help="full: full-N force pass at N=2^20 (default); blocks: i-blocks of "
                         "2^k against N=65536 with per-j times, FP64 vs mixed errors")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the force path is built and measured on a single "
              "MI355X (multi-GPU j-domain sharding is a design note only)", file=sys.stderr)
        return 2
    if args.copy_calib:
        return copy_calibration()
    if args.workload == "blocks":
        return blocks_study(args)
    if args.workload == "tiles":
        return tile_sweep(args)
    if args.workload == "ph4":
        return ph4_c0(args)
    if args.workload == "clusters":
        return clusters_c2(args)
    if args.workload == "adjoint":
        return adjoint_c3(args)
    if args.workload == "driver":
        return driver_study(args)
    if args.workload == "worker":
        return worker_study(args)
    if args.workload == "c4evolve":
        return c4_evolve(args)

    from amuse_amd.g6 import G6
    n, eps2 = args.n, args.eps2
    log(f"bench: building N={n} Plummer/Salpeter system")
    m, pos, vel = make_system(n)
    g6 = G6(0)
    g6.open()
    upload(g6, m, pos, vel)
    pairs = float(n) * float(n)

    log("bench: FP64 path")
    w64, t64, r64 = time_path(g6, 0, n, eps2, args.steps, args.warmup)
    log("bench: mixed path")
    wmx, tmx, rmx = time_path(g6, 1, n, eps2, args.steps, args.warmup)
    g6.set_precision(0)

    gp64 = pairs * args.steps / w64 / 1e9
    gpmx = pairs * args.steps / wmx / 1e9
    k64 = t64["force_ms"] / max(t64["force_launches"], 1)
    kmx = tmx["force_ms"] / max(tmx["force_launches"], 1)
    pairs_per_launch_mx = tmx["pairs"] / max(tmx["force_launches"], 1)
    pairs_per_launch_64 = t64["pairs"] / max(t64["force_launches"], 1)
    ach_mx = FLOP_PER_PAIR * pairs_per_launch_mx / (kmx * 1e-3) / 1e12
    ach_64 = FLOP_PER_PAIR * pairs_per_launch_64 / (k64 * 1e-3) / 1e12

    # mixed vs FP64 error over all particles; potential energy from each path
    a64, j64, p64, nn64 = r64
    amx, jmx, pmx, nnmx = rmx
    with np.errstate(divide="ignore", invalid="ignore"):
        ea = np.linalg.norm(amx - a64, axis=1) / np.linalg.norm(a64, axis=1)
        ep = np.abs(pmx - p64) / np.abs(p64)
    U64 = 0.5 * float((m * p64).sum())
    Umx = 0.5 * float((m * pmx).sum())
    K = 0.5 * float((m * (vel * vel).sum(1)).sum())

    cpu = cpu_best = None
    host = host_info()
    if not args.no_cpu:
        threads = args.cpu_threads or usable_cores(host)
        log(f"bench: CPU baseline (g6lib restatement, {threads} threads; host {host})")
        sec, ni_cpu, cres = cpu_baseline(m, pos, vel, eps2, threads, args.cpu_per_thread)
        cpu_gps = ni_cpu * float(n) / sec / 1e9
        # the CPU sample's forces agree with the GPU FP64 path
        dc = np.linalg.norm(cres[0] - a64[:ni_cpu], axis=1) / np.linalg.norm(a64[:ni_cpu], axis=1)
        cpu = {"value": cpu_gps, "unit": "G pair-interactions/s", "cores": threads,
               "kind": "port",
               "sample": f"{ni_cpu} i-particles x {n} j (g6lib.c restatement: AoS records, "
                         f"pow() triple, per-call predictor, {args.cpu_per_thread} i per call), "
                         f"{threads} independent units on {threads} threads, {sec:.1f} s",
               "host": host,
               "per_core": cpu_gps / threads,
               "max_rel_acc_vs_gpu_fp64": float(dc.max())}
        # best-effort CPU row: ph4's own CPU pair loop, SIMD over j, threads over i
        ni_b = 128 * threads
        log(f"bench: best-effort CPU loop ({threads} threads, {ni_b} i)")
        import oracle
        bsec, ba, _, _, _ = oracle.cpu_best_run(threads, pos[:ni_b], vel[:ni_b], eps2, m, pos, vel)
        db = np.linalg.norm(ba - a64[:ni_b], axis=1) / np.linalg.norm(a64[:ni_b], axis=1)
        cpu_best = {"value": ni_b * float(n) / bsec / 1e9, "unit": "G pair-interactions/s",
                    "cores": threads, "kind": "port",
                    "sample": f"{ni_b} i-particles x {n} j, ph4's CPU pair loop "
                              f"(src/amuse_ph4/src/idata.cc:195-236) in FP64, AVX2 SIMD over j, "
                              f"OpenMP over i, {bsec:.1f} s",
                    "per_core": ni_b * float(n) / bsec / 1e9 / threads,
                    "max_rel_acc_vs_gpu_fp64": float(db.max())}

    g6.close()
    # c4's 8 block steps after the full-N evaluation, both paths, energies
    # from FP64 passes (the ph4-style driver over the same library)
    log("bench: 8 block steps, FP64 then mixed forces")
    evo = {"fp64": evolve_blocks(m, pos, vel, eps2, args.eta, 8, 0),
           "mixed": evolve_blocks(m, pos, vel, eps2, args.eta, 8, 1),
           "fp64_device_driver": evolve_blocks(m, pos, vel, eps2, args.eta, 8, 0, True),
           "mixed_device_driver": evolve_blocks(m, pos, vel, eps2, args.eta, 8, 1, True)}
    # the discriminating energy evidence: the same system integrated to
    # t = --c4-t (about 10^7 particle steps) with each path's forces
    evolve = None
    if args.c4_t > 0:
        log(f"bench: integration to t={args.c4_t}, FP64 then mixed forces")
        evolve = c4_evolve_run(m, pos, vel, eps2, args.eta, args.c4_t)

    traffic = load_traffic("mixed_full_n")
    out = {
        "metric": "G pair-interactions/s (acc+jerk+pot) at N=1M, mixed vs FP64 at stated "
                  "energy error",
        "value": gpmx,
        "unit": "G pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wmx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "ds+fp32",
        "data": "synthetic",
        "config": {"workload": f"full-N force evaluation (acc, jerk, pot, nn), N={n} Plummer "
                               f"sphere, Salpeter IMF 0.1-100 normalised to M=1, eps2={eps2}",
                   "n": n, "pairs_per_step": pairs, "eps2": eps2},
        "fp64": {"value": gp64, "ms_per_step": w64 / args.steps * 1e3, "dtype": "f64",
                 "kernel_ms": k64, "achieved_tflops": ach_64,
                 "frac_fp64_peak": ach_64 / PEAK_TFLOPS["fp64"]},
        "mixed_over_fp64": gpmx / gp64,
        "pairs_excluding_self_per_step": float(n) * float(n - 1),
        "accuracy": {"mixed_vs_fp64_acc_rel_median": float(np.median(ea)),
                     "mixed_vs_fp64_acc_rel_p99": float(np.quantile(ea, 0.99)),
                     "mixed_vs_fp64_pot_rel_max": float(ep.max()),
                     "U_fp64": U64, "U_mixed": Umx, "K": K,
                     "energy_rel_diff_mixed_vs_fp64": abs(Umx - U64) / abs(K + U64),
                     "nn_agreement": float((nnmx == nn64).mean())},
        "roofline": {"bound": "valu", "achieved": ach_mx, "peak": PEAK_TFLOPS["mixed"],
                     "unit": "TFLOP/s", "frac": ach_mx / PEAK_TFLOPS["mixed"],
                     "traffic": traffic, "kernel_ms": kmx,
                     "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": pairs_per_launch_mx},
        "cpu_baseline": cpu,
        "cpu_best_effort": cpu_best,
        "block_steps_8": evo,
        "energy_error": None if evolve is None else {
            "fp64": evolve["fp64"]["dE_over_E"], "mixed": evolve["mixed"]["dE_over_E"],
            "ratio_mixed_over_fp64": evolve["energy_error_ratio_mixed_over_fp64"],
            "eta": args.eta, "t_end": args.c4_t,
            "particle_steps": {"fp64": evolve["fp64"]["particle_steps"],
                               "mixed": evolve["mixed"]["particle_steps"]},
            "block_steps": {"fp64": evolve["fp64"]["block_steps"],
                            "mixed": evolve["mixed"]["block_steps"]},
            "gpairs_per_s": {"fp64": evolve["fp64"]["gpairs_per_s"],
                             "mixed": evolve["mixed"]["gpairs_per_s"]},
            "note": "relative energy change of the same N=2^20 system integrated from t=0 "
                    "with ph4's block steps on the device-resident driver, step forces in "
                    "the named path (initial forces and steps in FP64 for both), energies "
                    "from FP64 passes"},
        "energy_error_8_block_steps": {"fp64": evo["fp64"]["dE_over_E"],
                                       "mixed": evo["mixed"]["dE_over_E"]},
    }
    if cpu:
        out["speedup_vs_cpu"] = {"mixed": gpmx / cpu["value"], "fp64": gp64 / cpu["value"]}
        out["speedup_vs_cpu_best_effort"] = {"mixed": gpmx / cpu_best["value"],
                                             "fp64": gp64 / cpu_best["value"]}
        # every logical CPU of the host at the measured per-core rate: an upper
        # bound on what the host could reach beyond the cores this run may use
        nproc = host["nproc"] or cpu["cores"]
        out["speedup_vs_cpu_best_effort_all_host_cpus"] = {
            "cpus": nproc, "cpu_value": cpu_best["per_core"] * nproc,
            "mixed": gpmx / (cpu_best["per_core"] * nproc),
            "fp64": gp64 / (cpu_best["per_core"] * nproc)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
