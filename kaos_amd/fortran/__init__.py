"""Fortran driver of libkaos_hip.so (kaos_fdriver) and its file formats.

kaos_hip_iface.f90 is the iso_c_binding module a Fortran caller uses;
kaos_fdriver.f90 drives the hot path through it.  write_input / read_output
produce and parse the driver's stream files (layout in kaos_fdriver.f90)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "kaos_fdriver")
CONFIG_ARRAYS = ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Te")


def write_input(path, tube, q0, vpar0, vperp0, nsteps, ndatfac=1, precision=0, capacity=None, seed=1,
                export=None):
    """export: None, or dict(theta, Ts, density (nq each), phi, ns0) for a run
    with an export directory (and optionally device-seeded ions)."""
    a = tube.arrays
    n = len(q0)
    cap = capacity if capacity is not None else n + (nsteps + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 1024
    with open(path, "wb") as f:
        np.array([tube.nq, tube.nvpar, tube.nvperp, tube.grav, tube.eamb, tube.eambsign, tube.filter_pt,
                  tube.lbcond, tube.ubcond, ndatfac, precision, nsteps], dtype=np.int32).tofile(f)
        np.array([tube.n_inj_lb, tube.n_inj_ub, n, cap, seed], dtype=np.int64).tofile(f)
        np.array([tube.p, tube.cphi, tube.sphi, tube.h, tube.ms, tube.qs, tube.nsnormfac, tube.Ts_lb,
                  tube.Ts_ub], dtype=np.float64).tofile(f)
        for k in CONFIG_ARRAYS:
            np.ascontiguousarray(a[k], dtype=np.float64).tofile(f)
        for k in ("vparGL", "vparGH", "vparGC", "dvpar", "vperpGL", "vperpGH", "vperpGC", "dvperp", "d3v"):
            np.ascontiguousarray(a[k], dtype=np.float64).tofile(f)
        for v in (q0, vpar0, vperp0):
            np.ascontiguousarray(v, dtype=np.float64).tofile(f)
        if export is not None:
            for k in ("theta", "Ts", "density"):
                np.ascontiguousarray(export[k], dtype=np.float64).tofile(f)
            np.array([export["phi"], export["ns0"]], dtype=np.float64).tofile(f)


def write_grid3(path, tube):
    """The driver's GRID3D file of a tube with a 3-D velocity grid."""
    a = tube.arrays
    with open(path, "wb") as f:
        np.array([tube.nv1, tube.nv2], dtype=np.int32).tofile(f)
        for k in ("v1GL", "v1GH", "v1GC", "dv1", "v2GL", "v2GH", "v2GC", "dv2", "d3v3"):
            np.ascontiguousarray(a[k], dtype=np.float64).tofile(f)


def read_output(path, nq, v3d=False):
    """List of master-step records: dict(n, n_alive, nq_counts, M (6, nq), EA, esc[, M8 (8, nq)])."""
    fields = [("n", "<i8"), ("alive", "<i8"), ("nqc", "<i8", (nq,)), ("M", "<f8", (6 * nq,)),
              ("EA", "<f8", (nq,)), ("esc", "<f8", (4,))]
    if v3d:
        fields.append(("M8", "<f8", (8 * nq,)))
    raw = np.fromfile(path, dtype=np.dtype(fields))
    out = []
    for r in raw:
        d = dict(n=int(r["n"]), n_alive=int(r["alive"]), nq_counts=r["nqc"].copy(),
                 M=r["M"].reshape(6, nq).copy(), EA=r["EA"].copy(), esc=r["esc"].copy())
        if v3d:
            d["M8"] = r["M8"].reshape(8, nq).copy()
        out.append(d)
    return out
