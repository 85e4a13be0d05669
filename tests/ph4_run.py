"""One run of a block-step Hermite integrator on a GRAPE-6 library, printed as
JSON; run in its own process because each library holds one global unit.

    python tests/ph4_run.py --integrator ph4|own --lib PATH --n N --t T [--seed S]
    python tests/ph4_run.py --integrator ph4|own --lib PATH --ic line7 \\
        --eps2 1e-10 --eta 0.14 --touts 0,0.23,0.24,0.26,0.27

ph4     the reference ph4 integrator (built unchanged by `make -C oracle ph4run`
        into oracle/_ref/libph4run_<backend>.so, driven by oracle/ph4_harness.cc)
own     amuse_amd/lib/libhermite4.so bound at run time to --lib
device  the device-resident driver of amuse_amd/lib/libg6.so (--lib unused)

--sync: every particle synchronised to the system time after each output
(ph4's synchronize_all) before the energies are taken (ph4 and device).

Initial conditions: a Plummer sphere (amuse_amd.ic.plummer, --n, --seed), or
`line7`, the seven bodies on a line of src/amuse_ph4/tests/test_ph4.py:973-981
(m = 0.1; x = -101, -100, -2, 2, 98, 102, 104; vx = 2, -2, 2, -2, 2, -2, -5),
whose first two bodies meet head on at t = 1/4.  Output times: --touts, or
--nout + 1 equal intervals of [0, --t].  Each output holds the energies, the
block and particle step counts and the smallest particle time step.
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from amuse_amd import ic  # noqa: E402


def line7():
    m = np.full(7, 0.1)
    x = np.zeros((7, 3))
    x[:, 0] = [-101.0, -100.0, -2.0, 2.0, 98.0, 102.0, 104.0]
    v = np.zeros((7, 3))
    v[:, 0] = [2.0, -2.0, 2.0, -2.0, 2.0, -2.0, -5.0]
    return m, x, v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--integrator", choices=["ph4", "own", "device"], required=True)
    ap.add_argument("--sync", action="store_true")
    ap.add_argument("--lib", required=True)
    ap.add_argument("--ic", choices=["plummer", "line7"], default="plummer")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--t", type=float, default=0.125)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--eps2", type=float, default=1e-4)
    ap.add_argument("--eta", type=float, default=0.01)
    ap.add_argument("--nout", type=int, default=4)
    ap.add_argument("--touts", default="", help="comma-separated output times")
    a = ap.parse_args()
    m, x, v = line7() if a.ic == "line7" else ic.plummer(a.n, a.seed)
    n = m.size
    if a.touts:
        touts = np.array([float(q) for q in a.touts.split(",")])
    else:
        touts = a.t * np.arange(0, a.nout + 1) / a.nout   # includes t = 0
    out = {"t": touts.tolist()}
    if a.integrator == "ph4":
        lib = C.CDLL(a.lib)
        D, L = C.POINTER(C.c_double), C.POINTER(C.c_longlong)
        nt = touts.size
        k, p, dtmin = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        b, s = np.zeros(nt, np.int64), np.zeros(nt, np.int64)
        rc = lib.ph4h_run_sync(n, m.ctypes.data_as(D), x.ctypes.data_as(D),
                               v.ctypes.data_as(D), C.c_double(a.eps2), C.c_double(a.eta), nt,
                               touts.ctypes.data_as(D), k.ctypes.data_as(D),
                               p.ctypes.data_as(D), b.ctypes.data_as(L), s.ctypes.data_as(L),
                               dtmin.ctypes.data_as(D), int(a.sync))
        out.update(E=(k + p).tolist(), K=k.tolist(), blocks=b.tolist(), steps=s.tolist(),
                   dtmin=dtmin.tolist(), used_library=rc)
    elif a.integrator == "device":
        from amuse_amd.blocks import DeviceHermite4
        h = DeviceHermite4(precision=0)
        h.commit(m, x, v, eps2=a.eps2, eta=a.eta)
        E, K, B, S, DT = [], [], [], [], []
        for t in touts:
            h.evolve(t)
            if a.sync:
                h.synchronize()
            kk, uu = h.energy()
            c = h.counters()
            E.append(kk + uu); K.append(kk)
            B.append(c["block_steps"]); S.append(c["particle_steps"])
            DT.append(float(h.timesteps()[1].min()))
        h.close()
        out.update(E=E, K=K, blocks=B, steps=S, dtmin=DT)
    else:
        from amuse_amd.hermite import Hermite4
        h = Hermite4(a.lib)
        h.commit(m, x, v, eps2=a.eps2, eta=a.eta)
        E, K, B, S, DT = [], [], [], [], []
        for t in touts:
            h.evolve(t)
            kk, uu = h.energy()
            c = h.counters()
            E.append(kk + uu); K.append(kk)
            B.append(c["block_steps"]); S.append(c["particle_steps"])
            DT.append(float(h.timesteps()[1].min()))
        h.close()
        out.update(E=E, K=K, blocks=B, steps=S, dtmin=DT)
    print("RESULT " + json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
