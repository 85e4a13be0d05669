"""Golden run of the seven bodies on a line of src/amuse_ph4/tests/
test_ph4.py:970-1004 (eps2 = 1e-10, ph4's default eta = 0.14; bodies 1 and
2 meet head on at t = 1/4) through the reference ph4 integrator
(src/amuse_ph4/src built unchanged with -DGPU by `make -C oracle ph4run`) on
the CPU oracle behind the GRAPE-6 ABI, with output at t = 0, 0.23, 0.24,
0.26, 0.27 as in that test:

    python tests/golden/make_ph4_line7.py      (where the reference checkout
                                                is present; a few seconds)

Writes tests/golden/ph4_line7.npz (times, K, E, block and particle step
counts, smallest particle time step).
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TOUTS = "0,0.23,0.24,0.26,0.27"


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libph4run_oracle.so")
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ph4_run.py"),
                          "--integrator", "ph4", "--lib", lib, "--ic", "line7", "--eps2", "1e-10",
                          "--eta", "0.14", "--touts", TOUTS], capture_output=True, text=True,
                         check=True)
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    r = json.loads(line[7:])
    assert r["used_library"] == 1
    np.savez_compressed(os.path.join(HERE, "ph4_line7.npz"), t=np.array(r["t"]),
                        E=np.array(r["E"]), K=np.array(r["K"]), blocks=np.array(r["blocks"]),
                        steps=np.array(r["steps"]), dtmin=np.array(r["dtmin"]), eps2=1e-10,
                        eta=0.14)
    print(r)


if __name__ == "__main__":
    main()
