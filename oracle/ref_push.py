"""TEST INFRASTRUCTURE ONLY -- run the reference's own compiled push.

oracle/_ref/kaos_ref_push is built by oracle/ref/Makefile from the reference
Fortran sources (KineticRK4UpdateSub and its callees, unmodified) and a
harness that sets the module globals the push reads.  This module writes the
ions of a tube to the harness's input format, runs it and reads back
xN..VzN, Vpar, qk4 and the cell indices.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "kaos_ref_push")
REF_DIR = "/root/reference"


def available():
    return os.path.exists(EXE)


def build():
    """Build the harness when the reference sources are present (not on the GPU box)."""
    if not os.path.isdir(REF_DIR):
        return None
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref")], check=True)
    return EXE


def run(tube, x, y, z, Vx, Vy, Vz, vperp, aea, ag, qind, vparind, vperpind, n=5, ndatfac=1000):
    ns = len(x)
    a = tube.arrays
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([tube.nq, tube.nvpar, tube.nvperp, ns], np.int64).tofile(f)
            np.array([tube.grav, tube.eamb, tube.eambsign, 0, 0, n, ndatfac, 0], np.int64).tofile(f)
            np.array([tube.p, tube.phi, tube.h, tube.ms, tube.qs, 0, 0, 0], np.float64).tofile(f)
            for k in ("qGL", "qGH", "vparGL", "vparGH", "vperpGL", "vperpGH"):
                np.asarray(a[k], np.float64).tofile(f)
            for arr in (x, y, z, Vx, Vy, Vz, vperp, aea, ag):
                np.asarray(arr, np.float64).tofile(f)
            for arr in (qind, vparind, vperpind):
                np.asarray(arr, np.int64).tofile(f)
        r = subprocess.run([EXE, fin, fout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        raw = np.fromfile(fout, dtype=np.float64, count=8 * ns)
        ints = np.fromfile(fout, dtype=np.int64, offset=8 * 8 * ns)
    names = ("xN", "yN", "zN", "VxN", "VyN", "VzN", "Vpar", "qk4")
    out = {k: raw[i * ns:(i + 1) * ns].copy() for i, k in enumerate(names)}
    out.update(qind=ints[:ns].copy(), vparind=ints[ns:2 * ns].copy(), vperpind=ints[2 * ns:].copy())
    out["stdout"] = r.stdout
    for line in r.stdout.splitlines():
        if line.strip().startswith("push_seconds"):
            out["push_seconds"] = float(line.split()[1])
    return out
