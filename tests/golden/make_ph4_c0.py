"""Golden energy trajectory of configuration 0 from the reference ph4
integrator (src/amuse_ph4/src built unchanged with -DGPU by `make -C oracle
ph4run`) running on the CPU oracle behind the GRAPE-6 ABI:

    Plummer N = 1024 (amuse_amd.ic.plummer, seed 1024), eps2 = 1e-4,
    eta = 0.01, t = 0 .. 10, energies every 0.5.

    python tests/golden/make_ph4_c0.py      (this container, ~10 min)

Writes tests/golden/ph4_c0_energy.npz (times, K, E, block and particle
step counts).
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libph4run_oracle.so")
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ph4_run.py"),
                          "--integrator", "ph4", "--lib", lib, "--n", "1024", "--t", "10",
                          "--seed", "1024", "--nout", "20"], capture_output=True, text=True,
                         check=True)
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    r = json.loads(line[7:])
    np.savez_compressed(os.path.join(HERE, "ph4_c0_energy.npz"), t=np.array(r["t"]),
                        E=np.array(r["E"]), K=np.array(r["K"]), blocks=np.array(r["blocks"]),
                        steps=np.array(r["steps"]), n=1024, seed=1024, eps2=1e-4, eta=0.01)
    print(r)


if __name__ == "__main__":
    main()
