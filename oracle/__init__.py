"""TEST INFRASTRUCTURE ONLY -- the CPU oracle of the KAOS hot path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker; the shipped library never calls it.

liboracle.so is oracle/kaos_oracle.c built with gcc -O2 -ffp-contract=off
(see oracle/build.py).  The Oracle class mirrors kaos_amd.solver.KineticSolver
for one tube so the parity tests read the same on both sides.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)

ICR_FIELDS = ("icr_lambda_perp", "icr_eta_lh", "icr_xi_perp1", "icr_xi_perp2", "icr_s0", "icr_f0",
              "icr_chi_perp1", "icr_chi_perp2")
ARRAYS = ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Te",
          "vparGL", "vparGH", "vparGC", "dvpar", "vperpGL", "vperpGH", "vperpGC", "dvperp", "d3v")
MOMENTS3 = ("M0", "M1perp1", "M1perp2", "M1par", "M2", "M2perp1", "M2perp2", "M2par")
ARRAYS3 = ("v1GL", "v1GH", "v1GC", "dv1", "v2GL", "v2GH", "v2GC", "dv2", "d3v3")
ENA_ARRAYS = ("vpGL", "vpGH", "vpGC", "dvp", "vqGL", "vqGH", "vqGC", "dvq", "vphiGL", "vphiGH", "vphiGC", "dvphi",
              "ena_hh", "ena_d33v")
ENA_MOMENTS = ("M0", "M1P", "M1Q", "M1PHI", "M2", "M2P", "M2Q", "M2PHI")


class OrcCfg(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nq", "nvpar", "nvperp", "grav", "eamb", "eambsign",
                                        "filter_pt", "ndatfac", "lbcond", "ubcond")]
                + [("n_inj_lb", C.c_int64), ("n_inj_ub", C.c_int64)]
                + [(n, C.c_double) for n in ("p", "cphi", "sphi", "h", "ms", "qs", "nsnormfac",
                                             "Ts_lb", "Ts_ub")]
                + [("tube", C.c_int), ("seed", C.c_uint64)]
                + [(n, _dp) for n in ARRAYS]
                + [("icr", C.c_int)]
                + [(n, C.c_double) for n in ICR_FIELDS]
                + [("filter_pt_m1par", C.c_int)]
                + [("v3d", C.c_int), ("nv1", C.c_int), ("nv2", C.c_int)]
                + [(n, _dp) for n in ARRAYS3]
                + [(n, C.c_int) for n in ("cx", "cx_filter", "cx_filter_pt", "pad_cx")]
                + [("t_neut", C.c_double), ("m_neut", C.c_double), ("cx_sigma_scale", C.c_double),
                   ("nsneut", _dp), ("ena_cap", C.c_int64)]
                + [(n, C.c_int) for n in ("nvp", "nvq", "nvphi", "pad_ena")]
                + [(n, _dp) for n in ENA_ARRAYS])


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build()
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i32p = C.POINTER(C.c_int)
        lib.orc_create.argtypes = [C.POINTER(OrcCfg), C.c_int64, _dp, _dp, _dp, C.c_int64]
        lib.orc_create.restype = P
        lib.orc_destroy.argtypes = [P]
        lib.orc_step.argtypes = [P, C.c_int64, _dp]
        lib.orc_step.restype = C.c_int
        lib.orc_count.argtypes = [P]
        lib.orc_count.restype = C.c_int64
        lib.orc_get_particles.argtypes = [P] + [_dp] * 10 + [i32p] * 3
        lib.orc_get_master.argtypes = [P] + [_dp] * 8
        lib.orc_filter.argtypes = [_dp, _dp, C.c_int, C.c_int]
        lib.orc_ambipolar.argtypes = [C.POINTER(OrcCfg), _dp, _dp, _dp]
        lib.orc_gravity.argtypes = [C.POINTER(OrcCfg), _dp]
        lib.orc_scan_qcell.argtypes = [C.POINTER(OrcCfg), C.c_double]
        lib.orc_scan_qcell.restype = C.c_int
        lib.orc_scan_vcell.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int]
        lib.orc_scan_vcell.restype = C.c_int
        lib.orc_moments_from_counts.argtypes = [C.POINTER(OrcCfg), _dp, _dp, _dp, _dp]
        lib.orc_push_lines.argtypes = [C.POINTER(OrcCfg), C.c_int64, _dp, _dp, _dp, _dp, _dp]
        lib.orc_cartesian.argtypes = [C.POINTER(OrcCfg), C.c_int64, _dp, _dp, _dp]
        lib.orc_get_lines.argtypes = [P, _dp, _dp]
        lib.orc_find_qcell_from.argtypes = [C.POINTER(OrcCfg), C.c_double, C.c_int]
        lib.orc_find_qcell_from.restype = C.c_int
        lib.orc_find_vcell_guess.argtypes = [_dp, _dp, C.c_int, C.c_double]
        lib.orc_find_vcell_guess.restype = C.c_int
        lib.orc_interp2_from.argtypes = [C.POINTER(OrcCfg), _dp, C.c_double, C.c_int, _dp]
        lib.orc_interp2_from.restype = C.c_int
        for n in ("orc_root_r",):
            getattr(lib, n).argtypes = [C.c_double, C.c_double]
            getattr(lib, n).restype = C.c_double
        lib.orc_quartic_newton.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        lib.orc_quartic_newton.restype = C.c_double
        for n in ("orc_sinh_asinh_third", "orc_log", "orc_cos2pi"):
            getattr(lib, n).argtypes = [C.c_double]
            getattr(lib, n).restype = C.c_double
        lib.orc_ea_stencil.argtypes = [C.POINTER(OrcCfg), _dp, _dp, _dp]
        lib.orc_find_qcell.argtypes = [C.POINTER(OrcCfg), C.c_double]
        lib.orc_find_qcell.restype = C.c_int
        lib.orc_find_vcell.argtypes = [_dp, _dp, C.c_int, C.c_double]
        lib.orc_find_vcell.restype = C.c_int
        lib.orc_interp_cells.argtypes = [C.POINTER(OrcCfg), _dp, C.c_double, _dp]
        lib.orc_interp_cells.restype = C.c_int
        lib.orc_interp_scan.argtypes = [C.POINTER(OrcCfg), _dp, C.c_double, C.c_double, C.c_int]
        lib.orc_interp_scan.restype = C.c_double
        lib.orc_mb.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        lib.orc_gauss.argtypes = [C.c_double, C.c_double]
        lib.orc_gauss.restype = C.c_double
        u64p = C.POINTER(C.c_uint64)
        lib.orc_philox.argtypes = [u64p, u64p, u64p]
        lib.orc_inject_uniforms.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int64, _dp]
        lib.orc_exp.argtypes = [C.c_double]
        lib.orc_exp.restype = C.c_double
        lib.orc_sin2pi.argtypes = [C.c_double]
        lib.orc_sin2pi.restype = C.c_double
        lib.orc_pow_neg.argtypes = [C.c_double, C.c_double]
        lib.orc_pow_neg.restype = C.c_double
        u32p = C.POINTER(C.c_uint32)
        lib.orc_philox4x32.argtypes = [u32p, u32p, u32p]
        lib.orc_icr_uniforms.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int64, _dp]
        lib.orc_icr_kick.argtypes = [C.POINTER(OrcCfg), C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, _dp]
        lib.orc_set_vperp.argtypes = [P, _dp, _dp]
        lib.orc_set_q0.argtypes = [P, _dp]
        lib.orc_set_field_input.argtypes = [P, _dp, C.c_int]
        lib.orc_get_vperp12.argtypes = [P, _dp, _dp, C.POINTER(C.c_int64)]
        i64p = C.POINTER(C.c_int64)
        lib.orc_cell_count.argtypes = [C.c_double, C.c_double, C.c_double]
        lib.orc_cell_count.restype = C.c_int64
        lib.orc_seed_ions.argtypes = [C.POINTER(OrcCfg), i64p, _dp, C.c_uint64, C.c_int] + [_dp] * 6
        lib.orc_seed_ions.restype = C.c_int64
        lib.orc_seed_uniforms.argtypes = [C.c_uint64, C.c_int, C.c_int64, _dp]
        lib.orc_get_m1f.argtypes = [P, _dp]
        lib.orc_get_cx.argtypes = [P, _dp, _dp, _dp]
        lib.orc_ena_count.argtypes = [P]
        lib.orc_ena_count.restype = C.c_int64
        lib.orc_get_enas.argtypes = [P, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.orc_get_ena_master.argtypes = [P, _dp, _dp]
        lib.orc_get_ena_moments.argtypes = [P, _dp, _dp]
        lib.orc_ena_moments_from_counts.argtypes = [C.POINTER(OrcCfg), _dp, _dp, _dp]
        lib.orc_ena_vcomp.argtypes = [_dp, _dp]
        lib.orc_cx_cell.argtypes = [_dp, C.c_uint64, C.c_int64, C.c_int, C.c_int, _dp]
        lib.orc_poisson.argtypes = [C.c_double, C.c_uint64, C.c_int64, C.c_int, C.c_int]
        lib.orc_poisson.restype = C.c_double
        lib.orc_cx_sigma.argtypes = [C.c_double]
        lib.orc_cx_sigma.restype = C.c_double
        lib.orc_ena_from_ion.argtypes = [C.c_double] * 8 + [C.c_double, _dp]
        lib.orc_ena_push.argtypes = [_dp, C.c_double, C.c_int]
        lib.orc_cx_block.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_uint64, _dp]
        lib.orc_get_master3.argtypes = [P, _dp, _dp]
        lib.orc_generate_grid.argtypes = ([C.c_int64] + [C.c_double] * 8 + [C.c_int] * 3 + [C.c_int64] * 2
                                          + [C.c_double] * 5 + [_dp, i64p, _dp, _dp, _dp, _dp])
        lib.orc_gg_qpoint.argtypes = [C.c_double, C.c_double, C.c_int64, C.c_int64]
        lib.orc_gg_qpoint.restype = C.c_double
        lib.orc_sin.argtypes = [C.c_double]
        lib.orc_sin.restype = C.c_double
        lib.orc_gg_ena_metric.argtypes = [C.c_double] * 8 + [_dp]
        _lib = lib
    return _lib


def ptr(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def cfg_of(tube, ndatfac=1, seed=1, tube_index=0):
    c = OrcCfg()
    for n in ("nq", "nvpar", "nvperp", "grav", "eamb", "eambsign", "filter_pt", "lbcond", "ubcond",
              "n_inj_lb", "n_inj_ub", "p", "h", "ms", "qs", "nsnormfac", "Ts_lb", "Ts_ub"):
        setattr(c, n, getattr(tube, n))
    c.cphi, c.sphi = tube.cphi, tube.sphi
    c.ndatfac = ndatfac
    c.tube = tube_index
    c.seed = seed
    c.icr = getattr(tube, "icr", 0)
    c.filter_pt_m1par = getattr(tube, "filter_pt_m1par", 0)
    c.v3d = getattr(tube, "v3d", 0)
    if c.v3d:
        c.nv1, c.nv2 = tube.nv1, tube.nv2
    for n in ICR_FIELDS:
        setattr(c, n, getattr(tube, n, 0.0))
    keep = []
    c.cx = getattr(tube, "cx", 0)
    if c.cx:
        if not c.v3d:
            raise ValueError("charge exchange needs the 3-D velocity grid (M1Perp1, M1Perp2)")
        c.cx_filter, c.cx_filter_pt = tube.cx_filter, tube.cx_filter_pt
        c.t_neut, c.m_neut, c.ena_cap = tube.t_neut, tube.m_neut, tube.ena_cap
        c.cx_sigma_scale = tube.cx_sigma_scale
        a = np.ascontiguousarray(tube.arrays["nsneut"], dtype=np.float64)
        keep.append(a)
        c.nsneut = ptr(a)
        c.nvp, c.nvq, c.nvphi = tube.nvp, tube.nvq, tube.nvphi
        for n in ENA_ARRAYS:
            a = np.ascontiguousarray(tube.arrays[n], dtype=np.float64)
            keep.append(a)
            setattr(c, n, ptr(a))
    for n in ARRAYS + (ARRAYS3 if c.v3d else ()):
        a = np.ascontiguousarray(tube.arrays[n], dtype=np.float64)
        keep.append(a)
        setattr(c, n, ptr(a))
    c._keep = keep
    return c


class Oracle:
    """One flux tube stepped by the CPU restatement."""

    def __init__(self, tube, q0, vpar0, vperp0, ndatfac=1, seed=1, capacity=None, tube_index=0):
        self.lib = load()
        self.tube = tube
        self.cfg = cfg_of(tube, ndatfac, seed, tube_index)
        q0, vpar0, vperp0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (q0, vpar0, vperp0))
        cap = capacity if capacity is not None else len(q0) + 1024
        self.h = self.lib.orc_create(C.byref(self.cfg), len(q0), ptr(q0), ptr(vpar0), ptr(vperp0), cap)

    def step(self, n, uniforms=None):
        u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64)
        rc = self.lib.orc_step(self.h, n, ptr(u))
        if rc:
            raise RuntimeError(f"oracle step {n} failed: {rc}")

    def count(self):
        return self.lib.orc_count(self.h)

    def set_q0(self, q0):
        """Replace the q0 that step 1 clamps, scans and interpolates the first
        fields at (the positions stay those of the q0 given at creation)."""
        self.lib.orc_set_q0(self.h, ptr(np.ascontiguousarray(q0, dtype=np.float64)))

    def set_field_input(self, EA, mode):
        """Restart field (EAmagInputT, nq values): mode 1 replaces the computed
        EAmagRT at step 1, mode 2 at step 1 and every master step, 0 computes it
        (_8_7_AmbipolarEfield.f90:1375-1407)."""
        e = None if EA is None else np.ascontiguousarray(EA, dtype=np.float64)
        self.lib.orc_set_field_input(self.h, ptr(e), int(mode))

    def cx(self):
        """Collision counts of the last master step: sigmaIonNeutRT, nuIonNeutRT,
        nuIonNeutPoiT per cell."""
        nq = self.tube.nq
        sg, nu, poi = np.zeros(nq), np.zeros(nq), np.zeros(nq)
        self.lib.orc_get_cx(self.h, ptr(sg), ptr(nu), ptr(poi))
        return {"sigma": sg, "nu": nu, "poi": poi}

    def enas(self):
        """The ENAs in id order: x, y, z, Vx, Vy, Vz, id, qind."""
        n = self.lib.orc_ena_count(self.h)
        X = np.zeros(6 * max(n, 1))
        ids, q = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
        self.lib.orc_get_enas(self.h, ptr(X), ptr(ids, C.c_int64), ptr(q, C.c_int64))
        X = X[:6 * n].reshape(n, 6)
        out = {k: X[:, i].copy() for i, k in enumerate(("x", "y", "z", "Vx", "Vy", "Vz"))}
        out["id"], out["qind"] = ids[:n], q[:n]
        return out

    def ena_moments(self):
        """ENA phase-space counts [Q][Vp][Vq][Vphi] and the eight ENA moments of the
        last master step."""
        t = self.tube
        n = np.zeros(t.nq * t.nvp * t.nvq * t.nvphi)
        M = np.zeros(8 * t.nq)
        self.lib.orc_get_ena_moments(self.h, ptr(n), ptr(M))
        out = {"counts": n.astype(np.int64).reshape(t.nq, t.nvp, t.nvq, t.nvphi)}
        for k, name in enumerate(ENA_MOMENTS):
            out[name] = M[k * t.nq:(k + 1) * t.nq].copy()
        return out

    def ena_master(self):
        """NqReNormENAT of the last master step and the LB/UB ENA escapes
        (counts, fluxes)."""
        nq = np.zeros(self.tube.nq)
        esc = np.zeros(4)
        self.lib.orc_get_ena_master(self.h, ptr(nq), ptr(esc))
        return {"nq_ena": nq, "esc": esc}

    def set_vperp(self, v1, v2):
        v1, v2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (v1, v2))
        self.lib.orc_set_vperp(self.h, ptr(v1), ptr(v2))

    def vperp12(self):
        n = self.count()
        v1, v2, ids = np.zeros(n), np.zeros(n), np.zeros(n, np.int64)
        self.lib.orc_get_vperp12(self.h, ptr(v1), ptr(v2), ptr(ids, C.c_int64))
        return v1, v2, ids

    def particles(self):
        n = self.count()
        d = {k: np.zeros(n) for k in ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "qk4", "aea", "ag")}
        i = {k: np.zeros(n, np.int32) for k in ("qind", "vparind", "vperpind")}
        self.lib.orc_get_particles(self.h, *(ptr(d[k]) for k in ("x", "y", "z", "Vx", "Vy", "Vz",
                                                                   "vperp", "qk4", "aea", "ag")),
                                   *(ptr(i[k], C.c_int) for k in ("qind", "vparind", "vperpind")))
        d.update(i)
        return d

    def lines(self):
        n = self.count()
        rs, vs = np.zeros(n), np.zeros(n)
        self.lib.orc_get_lines(self.h, ptr(rs), ptr(vs))
        return rs, vs

    def master(self):
        t = self.tube
        nq, nc = t.nq, t.nq * t.nvperp * t.nvpar
        nqc, nph, NqRT = np.zeros(nq), np.zeros(nc), np.zeros(nq)
        M, EA, EG, M0F, esc = np.zeros(6 * nq), np.zeros(nq), np.zeros(nq), np.zeros(nq), np.zeros(4)
        self.lib.orc_get_master(self.h, ptr(nqc), ptr(nph), ptr(NqRT), ptr(M), ptr(EA), ptr(EG),
                                ptr(M0F), ptr(esc))
        out = dict(nq_counts=nqc.astype(np.int64), nph_counts=nph.astype(np.int64).reshape(nq, t.nvperp, t.nvpar),
                   NqRT=NqRT, EA=EA, EG=EG, M0F=M0F, esc=esc)
        for k, n in enumerate(("M0", "M1perp", "M1par", "M2", "M2perp", "M2par")):
            out[n] = M[k * nq:(k + 1) * nq].copy()
        M1F = np.zeros(nq)
        self.lib.orc_get_m1f(self.h, ptr(M1F))
        out["M1F"] = M1F
        return out

    def master3(self):
        """3-D outputs of the last master step: counts [Q][V1][V2][Vpar] and the
        8 moments (M0, M1Perp1, M1Perp2, M1Par, M2, M2Perp1, M2Perp2, M2Par)."""
        t = self.tube
        n3, M8 = np.zeros(t.nq * t.nv1 * t.nv2 * t.nvpar), np.zeros(8 * t.nq)
        self.lib.orc_get_master3(self.h, ptr(n3), ptr(M8))
        out = dict(n3_counts=n3.astype(np.int64).reshape(t.nq, t.nv1, t.nv2, t.nvpar))
        for k, n in enumerate(MOMENTS3):
            out[n] = M8[k * t.nq:(k + 1) * t.nq].copy()
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def seed_ions(tube, counts, temperature, seed, tube_index=0):
    """The oracle's initial ions of one tube (orc_seed_ions): dict of q0, rs,
    vpar, v1, v2, vperp, ions numbered cell by cell."""
    lib = load()
    cfg = cfg_of(tube, tube_index=tube_index)
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    temperature = np.ascontiguousarray(temperature, dtype=np.float64)
    n = int(counts.sum())
    out = {k: np.zeros(n) for k in ("q0", "rs", "vpar", "v1", "v2", "vperp")}
    got = lib.orc_seed_ions(C.byref(cfg), ptr(counts, C.c_int64), ptr(temperature), seed, tube_index,
                            *(ptr(out[k]) for k in ("q0", "rs", "vpar", "v1", "v2", "vperp")))
    assert got == n
    return out


GRIDGEN_CFG = ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Ts", "Te", "density", "ns_neut")
GRIDGEN_AXES = ("vpar", "vperp", "vperp1", "vperp2", "vp", "vq", "vphi")


def generate_grid(p):
    """The oracle's grid set-up of one tube (orc_generate_grid) from the
    model inputs `p` (the keys of kaos_gridgen_in): dict of the per-cell
    arrays, counts, each axis as (GL, GH, GC, d), d3v (2-D or 3-D) and the
    ENA metric arrays hh, d33v [Q][bin]."""
    lib = load()
    nq = int(p["nq_ub"])
    npar, nperp = int(p["nvpar_gp"]) - 2, int(p["nvperp_gp"]) - 2
    v3, cx = int(p.get("vperp3d", 1)), int(p.get("cx", 0))
    nc = [npar, 0 if v3 else nperp, nperp if v3 else 0, nperp if v3 else 0,
          nperp if cx else 0, npar if cx else 0, nperp if cx else 0]
    cfg = np.zeros(13 * nq)
    counts = np.zeros(nq, np.int64)
    vel = np.zeros(max(1, 4 * sum(nc)))
    d3v = np.zeros(max(1, nperp * nperp * npar if v3 else nperp * npar))
    nb = nc[4] * nc[5] * nc[6]
    hh = np.zeros(max(1, nq * nb))
    d33 = np.zeros(max(1, nq * nb))
    rc = lib.orc_generate_grid(nq, p["qGA"], p["qGB"], p["L"], p["ms"], p["ti"], p["te"], p["ns0"], p["nsnormfac"],
                               v3, int(p.get("symvpar", 1)), cx, int(p["nvpar_gp"]), int(p["nvperp_gp"]),
                               p["vpar_sigma_fac"], p["vperp_sigma_fac"], p.get("t_neut", 0.0),
                               p.get("zns0_neut", 0.0), p.get("ns0_neut", 0.0), ptr(cfg), ptr(counts, C.c_int64),
                               ptr(vel), ptr(d3v), ptr(hh), ptr(d33))
    assert rc == 0
    out = {k: cfg[i * nq:(i + 1) * nq].copy() for i, k in enumerate(GRIDGEN_CFG)}
    out["counts"] = counts
    off = 0
    for a, n in zip(GRIDGEN_AXES, nc):
        out[a] = tuple(vel[off + k * n:off + (k + 1) * n].copy() for k in range(4))
        off += 4 * n
    out["d3v3" if v3 else "d3v"] = d3v
    out["ena_hh"], out["ena_d33v"] = hh[:nq * nb], d33[:nq * nb]
    out["ncell"] = dict(zip(GRIDGEN_AXES, nc))
    return out
