"""Writes the golden data of the 200-step integration check from the
reference program itself (oracle/_ref/magic, compiled from the reference's
sources by oracle/Makefile): for each case of tests/integration_cases.py,
the kinetic and outer-core magnetic energy series the program logs
(e_kin.TAG, e_mag_oc.TAG: its own output files, kept as written) and the
fields of the restart file it writes at the end (checkpoint_end.TAG, full
precision), as tests/golden/integ_<case>.npz.  Run in the build container:

    make -C oracle full && python tests/golden/make_integration_golden.py [case ...]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from magic_amd import checkpoint  # noqa: E402
import integration_cases as ic  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
MAGIC = os.path.join(ROOT, "oracle", "_ref", "magic")


def run_case(name):
    with tempfile.TemporaryDirectory() as td:
        tag = ic.prepare(name, td)
        env = dict(os.environ, OMP_NUM_THREADS="8", OMP_STACKSIZE="1G")
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; %s input.nml" % MAGIC], cwd=td,
                           env=env, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout[-3000:] + r.stderr[-3000:])
            raise RuntimeError("reference run failed: " + name)
        out = {"e_kin": np.loadtxt(os.path.join(td, "e_kin." + tag))}
        if ic.CASES[name][3]:
            out["e_mag_oc"] = np.loadtxt(os.path.join(td, "e_mag_oc." + tag))
        h = checkpoint.read_checkpoint(os.path.join(td, "checkpoint_end." + tag))
        out["time"] = np.float64(h["time"])
        out["r"] = h["r"]
        for k in ic.FIELDS:
            if h.get(k) is not None:
                out[k] = h[k]
    path = os.path.join(OUT, "integ_%s.npz" % name)
    np.savez_compressed(path, **out)
    print("wrote %s (%d bytes): %d log rows, final time %.12g" % (
        path, os.path.getsize(path), out["e_kin"].shape[0], float(out["time"])))


if __name__ == "__main__":
    for n in sys.argv[1:] or list(ic.CASES):
        run_case(n)
