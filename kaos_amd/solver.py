"""Host driver over the C-ABI: the Python mirror of KineticSolverSub's use of
the hot path (_8_KineticSolver.f90:35-263).

    ks = KineticSolver([tube], precision=KAOS_FP64, ndatfac=1, seed=1)
    ks.upload(0, q0, vpar0, vperp0)          # InitialConditionsSub output
    for n in range(1, Nt + 1):
        ks.step(n)                           # one ComputationalTimeStepLoop pass
    out = ks.master(0)                       # counts, moments, fields, escapes

Every number comes from libkaos_hip.so on the GPU; this module only moves
host arrays across the C-ABI.
"""
import ctypes as C
import numpy as np

from . import _abi
from ._abi import (KaosCfg, KaosGrid, KaosStepOut, KaosExportInfo, ptr, KAOS_FP64, KAOS_MIXED,  # noqa: F401
                   KAOS_FLAG_NO_GRAPH, KAOS_FLAG_NO_AUTOTUNE, KAOS_FLAG_VPERP3D,
                   KAOS_FIELD_COMPUTED, KAOS_FIELD_INPUT_FIRST, KAOS_FIELD_INPUT_FIXED)

MOMENT_NAMES = ("M0", "M1perp", "M1par", "M2", "M2perp", "M2par")
MOMENTS3 = ("M0", "M1perp1", "M1perp2", "M1par", "M2", "M2perp1", "M2perp2", "M2par")


class KaosError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kaos error {code}: {msg}")
        self.code = code


def grid_struct(tube):
    """kaos_grid for a grids.Tube; keeps the numpy arrays alive on the struct."""
    g = KaosGrid()
    for n in ("nq", "nvpar", "nvperp", "grav", "eamb", "eambsign", "filter_pt", "lbcond",
              "ubcond", "n_inj_lb", "n_inj_ub", "p", "h", "ms", "qs", "nsnormfac", "Ts_lb",
              "Ts_ub"):
        setattr(g, n, getattr(tube, n))
    g.cphi, g.sphi = tube.cphi, tube.sphi
    g.filter_pt_m1par = getattr(tube, "filter_pt_m1par", 0)
    if getattr(tube, "v3d", 0):
        g.nvperp1, g.nvperp2 = tube.nv1, tube.nv2
    g.icr = getattr(tube, "icr", 0)
    for n in _abi.ICR_FIELDS:
        setattr(g, n, getattr(tube, n, 0.0))
    keep = []
    for n in _abi.GRID_ARRAYS:
        a = np.ascontiguousarray(tube.arrays[n], dtype=np.float64)
        keep.append(a)
        setattr(g, n, ptr(a))
    if getattr(tube, "cx", 0):
        for n in ("cx", "cx_filter", "cx_filter_pt", "t_neut", "m_neut", "cx_sigma_scale", "ena_cap"):
            setattr(g, n, getattr(tube, n))
        a = np.ascontiguousarray(tube.arrays["nsneut"], dtype=np.float64)
        keep.append(a)
        g.ns_neut = ptr(a)
        g.nvp, g.nvq, g.nvphi = tube.nvp, tube.nvq, tube.nvphi
        for n in _abi.ENA_ARRAYS:
            a = np.ascontiguousarray(tube.arrays[n], dtype=np.float64)
            keep.append(a)
            setattr(g, n, ptr(a))
    if getattr(tube, "v3d", 0):
        for n, tn in zip(_abi.GRID_ARRAYS3, _abi.TUBE_ARRAYS3):
            if tube.arrays.get(tn) is None:
                continue                      # a null pointer: the library rejects the grid
            a = np.ascontiguousarray(tube.arrays[tn], dtype=np.float64)
            keep.append(a)
            setattr(g, n, ptr(a))
    g._keep = keep
    return g


class KineticSolver:
    def __init__(self, tubes, precision=KAOS_FP64, ndatfac=1, seed=1, capacity=None, device=0,
                 reorder_every=0, flags=0):
        self.lib = _abi.load()
        self.tubes = list(tubes)
        self.ndatfac = ndatfac
        v3 = [getattr(t, "v3d", 0) for t in self.tubes]
        if any(v3):
            if not all(v3):
                raise ValueError("the 3-D velocity grid is a context-wide mode: give every tube one")
            flags |= KAOS_FLAG_VPERP3D
        cfg = KaosCfg(ntubes=len(self.tubes), precision=precision, ndatfac=ndatfac, device=device,
                      seed=seed, reorder_every=reorder_every, flags=flags)
        h = C.c_void_p()
        rc = self.lib.kaos_init(C.byref(cfg), C.byref(h))
        if rc:
            raise KaosError(rc, self.lib.kaos_last_error(None).decode())
        self.h = h
        self._grids = []
        for t, tube in enumerate(self.tubes):
            g = grid_struct(tube)
            self._grids.append(g)
            cap = capacity[t] if capacity is not None else int(tube.cell_counts.sum()) + 1024
            self._chk(self.lib.kaos_set_grid(self.h, t, C.byref(g), int(cap)))

    def _chk(self, rc):
        if rc:
            raise KaosError(rc, self.lib.kaos_last_error(self.h).decode())

    def upload(self, tube, q0, vpar0, vperp0):
        q0, vpar0, vperp0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (q0, vpar0, vperp0))
        self._chk(self.lib.kaos_upload_particles(self.h, tube, len(q0), ptr(q0), ptr(vpar0),
                                                 ptr(vperp0)))

    def upload_vperp(self, tube, vperp1, vperp2):
        v1, v2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (vperp1, vperp2))
        self._chk(self.lib.kaos_upload_vperp(self.h, tube, len(v1), ptr(v1), ptr(v2)))

    def seed(self, tube, counts, temperature, seed):
        """Initial ions generated on the device (kaos_seed_particles): counts[Q]
        ions at the centre of cell Q, Maxwellian of temperature[Q] kelvin."""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        temperature = np.ascontiguousarray(temperature, dtype=np.float64)
        if len(counts) != self.tubes[tube].nq or len(temperature) != self.tubes[tube].nq:
            raise ValueError("counts and temperature need one entry per config cell")
        self._chk(self.lib.kaos_seed_particles(self.h, tube, ptr(counts, C.c_int64), ptr(temperature), seed))

    def step(self, n, uniforms=None, wait=True, noalloc=False):
        """One time step (kaos_step; kaos_step_noalloc with noalloc=True)."""
        out = KaosStepOut() if wait else None
        u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64)
        nu = 0 if u is None else len(u)
        fn = self.lib.kaos_step_noalloc if noalloc else self.lib.kaos_step
        self._chk(fn(self.h, n, ptr(u), nu, C.byref(out) if out is not None else None))
        return out

    def run(self, n_first, nsteps):
        """Steps n_first..n_first+nsteps-1 with no host sync between them; returns
        the number of ion pushes (particle-steps)."""
        ps = C.c_int64(0)
        self._chk(self.lib.kaos_run(self.h, n_first, nsteps, C.byref(ps)))
        return ps.value

    def set_launch(self, tiles_per_group, hist_words, persistent):
        self._chk(self.lib.kaos_set_launch(self.h, tiles_per_group, hist_words, int(persistent)))

    def get_launch(self):
        """(tiles_per_group, hist_words, persistent, tuning) of the push kernels."""
        v = [C.c_int32(0) for _ in range(4)]
        self._chk(self.lib.kaos_get_launch(self.h, *(C.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def set_launch_block(self, block):
        """Threads per block of the push (256 or 512; ends the autotuner)."""
        self._chk(self.lib.kaos_set_launch_block(self.h, int(block)))

    def get_launch_block(self):
        """Threads per block the push runs with (256 or 512)."""
        v = C.c_int32(0)
        self._chk(self.lib.kaos_get_launch_block(self.h, C.byref(v)))
        return v.value

    def profile(self, enable=True):
        self._chk(self.lib.kaos_profile(self.h, int(enable)))

    def profile_read(self):
        """(push-kernel ms summed, push launches, kaos_run ms summed) since the last read."""
        pm, nl, rm = C.c_double(0), C.c_int64(0), C.c_double(0)
        self._chk(self.lib.kaos_profile_read(self.h, C.byref(pm), C.byref(nl), C.byref(rm)))
        return pm.value, nl.value, rm.value

    def count(self, tube):
        n = C.c_int64(0)
        self._chk(self.lib.kaos_count(self.h, tube, C.byref(n)))
        return n.value

    def master(self, tube, phase=True):
        """Outputs of the last master step; phase=False leaves out the
        phase-space counts (nph_counts is then None)."""
        t = self.tubes[tube]
        nq, nc = t.nq, t.nq * t.nvperp * t.nvpar
        nqc = np.zeros(nq, np.int64)
        nph = np.zeros(nc, np.int64) if phase else None
        NqRT = np.zeros(nq)
        M = np.zeros(6 * nq)
        EA = np.zeros(nq)
        EG = np.zeros(nq)
        esc = np.zeros(4)
        self._chk(self.lib.kaos_download_master(self.h, tube, ptr(nqc, C.c_int64), ptr(nph, C.c_int64),
                                                ptr(NqRT), ptr(M), ptr(EA), ptr(EG), ptr(esc)))
        out = dict(nq_counts=nqc, nph_counts=None if nph is None else nph.reshape(nq, t.nvperp, t.nvpar), NqRT=NqRT,
                   EA=EA, EG=EG, esc=esc)
        for i, n in enumerate(MOMENT_NAMES):
            out[n] = M[i * nq:(i + 1) * nq].copy()
        return out

    def export_spinup(self, tube, directory, nq_lb=1, nq_ub=None, nsnorm_clb=None, nsnorm_cub=None,
                      lb_replenish=0, ub_replenish=0, species=1, flux_tube=1):
        """kaos_export_spinup: the spin-up files of the last master step
        (3-D grid only); defaults: cells 1..nq, the tube's injection counts."""
        t = self.tubes[tube]
        info = _abi.KaosSpinupInfo(
            dir=str(directory).encode(), species=species, flux_tube=flux_tube, nq_lb=nq_lb,
            nq_ub=t.nq if nq_ub is None else nq_ub,
            nsnorm_clb=float(t.n_inj_lb if nsnorm_clb is None else nsnorm_clb),
            nsnorm_cub=float(t.n_inj_ub if nsnorm_cub is None else nsnorm_cub),
            lb_replenish=lb_replenish, ub_replenish=ub_replenish)
        self._chk(self.lib.kaos_export_spinup(self.h, tube, C.byref(info)))

    def cx(self, tube):
        """kaos_download_cx: sigmaIonNeutRT, nuIonNeutRT, nuIonNeutPoiT of the
        last master step."""
        nq = self.tubes[tube].nq
        sg, nu, poi = np.zeros(nq), np.zeros(nq), np.zeros(nq)
        self._chk(self.lib.kaos_download_cx(self.h, tube, ptr(sg), ptr(nu), ptr(poi)))
        return {"sigma": sg, "nu": nu, "poi": poi}

    def enas(self, tube, n_cap=None):
        """kaos_download_enas: the tube's ENAs in id order."""
        n_cap = int(self.tubes[tube].ena_cap) if n_cap is None else n_cap
        bufs = [np.zeros(max(n_cap, 1)) for _ in range(6)]
        ids, q = np.zeros(max(n_cap, 1), np.int64), np.zeros(max(n_cap, 1), np.int64)
        n = C.c_int64(0)
        self._chk(self.lib.kaos_download_enas(self.h, tube, n_cap, C.byref(n), *[ptr(b) for b in bufs],
                                              ptr(ids, C.c_int64), ptr(q, C.c_int64)))
        k = n.value
        out = {name: b[:k].copy() for name, b in zip(("x", "y", "z", "Vx", "Vy", "Vz"), bufs)}
        out["id"], out["qind"] = ids[:k].copy(), q[:k].copy()
        return out

    def ena_moments(self, tube):
        """kaos_download_ena_moments: ENA phase-space counts [Q][Vp][Vq][Vphi] and
        M0, M1P, M1Q, M1PHI, M2, M2P, M2Q, M2PHI per cell."""
        t = self.tubes[tube]
        n = np.zeros(t.nq * t.nvp * t.nvq * t.nvphi, np.int64)
        M = np.zeros(8 * t.nq)
        self._chk(self.lib.kaos_download_ena_moments(self.h, tube, ptr(n, C.c_int64), ptr(M)))
        out = {"counts": n.reshape(t.nq, t.nvp, t.nvq, t.nvphi)}
        for k, name in enumerate(("M0", "M1P", "M1Q", "M1PHI", "M2", "M2P", "M2Q", "M2PHI")):
            out[name] = M[k * t.nq:(k + 1) * t.nq].copy()
        return out

    def ena_master(self, tube):
        """kaos_download_ena_master: NqReNormENAT and the LB/UB ENA escapes
        (counts, fluxes) of the last master step."""
        nq = np.zeros(self.tubes[tube].nq, np.int64)
        esc = np.zeros(4)
        self._chk(self.lib.kaos_download_ena_master(self.h, tube, ptr(nq, C.c_int64), ptr(esc)))
        return {"nq_ena": nq, "esc": esc}

    def set_field(self, tube, EA, mode):
        """kaos_set_field: input field EAmagInputT of a restart (nq values) and
        its mode (KAOS_FIELD_INPUT_FIRST, KAOS_FIELD_INPUT_FIXED or
        KAOS_FIELD_COMPUTED)."""
        e = None if EA is None else np.ascontiguousarray(EA, dtype=np.float64)
        self._chk(self.lib.kaos_set_field(self.h, tube, ptr(e), int(mode)))

    def master3(self, tube):
        """3-D outputs of the last master step (KAOS_FLAG_VPERP3D): counts
        [Q][V1][V2][Vpar] and the eight 3-D moments."""
        t = self.tubes[tube]
        n3 = np.zeros(t.nq * t.nv1 * t.nv2 * t.nvpar, np.int64)
        M = np.zeros(8 * t.nq)
        self._chk(self.lib.kaos_download_master3d(self.h, tube, ptr(n3, C.c_int64), ptr(M)))
        out = dict(n3_counts=n3.reshape(t.nq, t.nv1, t.nv2, t.nvpar))
        for i, n in enumerate(MOMENTS3):
            out[n] = M[i * t.nq:(i + 1) * t.nq].copy()
        return out

    def export_master(self, tube, directory, theta, Ts, ns0, phi, species=1, flux_tube=1):
        """DataExport2Sub's files of the last master step (kaos_export_master);
        returns the master-step index nn of the file names."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        Ts = np.ascontiguousarray(Ts, dtype=np.float64)
        if len(theta) != self.tubes[tube].nq or len(Ts) != self.tubes[tube].nq:
            raise ValueError("theta and Ts need one value per config cell")
        info = KaosExportInfo(dir=str(directory).encode(), species=species, flux_tube=flux_tube, ns0=ns0,
                              phi=phi, theta=ptr(theta), Ts=ptr(Ts))
        nn = C.c_int64(0)
        self._chk(self.lib.kaos_export_master(self.h, tube, C.byref(info), C.byref(nn)))
        return nn.value

    def particles(self, tube, cap=None, vperp12=False):
        """Live ions of a tube in the reference's order (kaos_download_particles;
        with vperp12 also Vperp1, Vperp2 from kaos_download_vperp)."""
        cap = int(cap if cap is not None else self.count(tube))
        d = {n: np.zeros(cap) for n in ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "qk4", "aea", "ag")}
        i = {n: np.zeros(cap, np.int64) for n in ("qind", "vparind", "vperpind")}
        n = C.c_int64(0)
        self._chk(self.lib.kaos_download_particles(
            self.h, tube, cap, C.byref(n), *(ptr(d[k]) for k in ("x", "y", "z", "Vx", "Vy", "Vz",
                                                                  "vperp", "qk4", "aea", "ag")),
            *(ptr(i[k], C.c_int64) for k in ("qind", "vparind", "vperpind"))))
        m = n.value
        out = {k: v[:m] for k, v in d.items()}
        out.update({k: v[:m] for k, v in i.items()})
        if vperp12:
            v1, v2 = np.zeros(cap), np.zeros(cap)
            self._chk(self.lib.kaos_download_vperp(self.h, tube, cap, C.byref(n), ptr(v1), ptr(v2)))
            out["vperp1"], out["vperp2"] = v1[:n.value], v2[:n.value]
        return out

    def upload_state(self, tube, P):
        """kaos_upload_state from a dict with the keys of particles(vperp12=True)
        (vperp1/vperp2 optional)."""
        d = {k: np.ascontiguousarray(P[k], dtype=np.float64)
             for k in ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qk4")}
        v1 = P.get("vperp1")
        v2 = P.get("vperp2")
        v1 = None if v1 is None else np.ascontiguousarray(v1, dtype=np.float64)
        v2 = None if v2 is None else np.ascontiguousarray(v2, dtype=np.float64)
        ii = {k: np.ascontiguousarray(P[k], dtype=np.int64) for k in ("qind", "vparind", "vperpind")}
        self._chk(self.lib.kaos_upload_state(
            self.h, tube, len(d["x"]), *(ptr(d[k]) for k in ("x", "y", "z", "Vx", "Vy", "Vz")), ptr(v1), ptr(v2),
            *(ptr(d[k]) for k in ("vperp", "aea", "ag", "qk4")),
            *(ptr(ii[k], C.c_int64) for k in ("qind", "vparind", "vperpind"))))

    def upload_moments(self, tube, M6):
        """kaos_upload_moments: the 6*nq moments of the last master step (the
        order of master()'s M0, M1perp, M1par, M2, M2perp, M2par)."""
        if isinstance(M6, dict):
            M6 = np.concatenate([M6[k] for k in MOMENT_NAMES])
        M6 = np.ascontiguousarray(M6, dtype=np.float64)
        if len(M6) != 6 * self.tubes[tube].nq:
            raise ValueError("6 moments per config cell")
        self._chk(self.lib.kaos_upload_moments(self.h, tube, ptr(M6)))

    def resume(self, n_last):
        self._chk(self.lib.kaos_resume(self.h, int(n_last)))

    def save_state(self, path):
        self._chk(self.lib.kaos_save_state(self.h, str(path).encode()))

    def load_state(self, path):
        self._chk(self.lib.kaos_load_state(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.kaos_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cell_counts(density, d3x, nsnormfac):
    """Macro-ions per config cell, (density / nsnormfac) d3x truncated toward zero
    (kaos_cell_counts; _0_1_1_ConfigGridGeneratorB.f90:83-85,
    _2_1_DensityProfileA1.f90:45).  Host arithmetic of the library: no GPU."""
    lib = _abi.load()
    density = np.ascontiguousarray(density, dtype=np.float64)
    d3x = np.ascontiguousarray(d3x, dtype=np.float64)
    if density.shape != d3x.shape or density.ndim != 1:
        raise ValueError("density and d3x must be 1-D arrays of one length")
    out = np.zeros(len(density), np.int64)
    rc = lib.kaos_cell_counts(len(density), ptr(density), ptr(d3x), float(nsnormfac), ptr(out, C.c_int64))
    if rc:
        raise KaosError(rc, "kaos_cell_counts: bad input (nsnormfac <= 0, or a count not finite)")
    return out


GRIDGEN_KEYS = ("nq_ub", "qGA", "qGB", "L", "phi", "ms", "ti", "te", "ns0", "nsnormfac", "vperp3d", "symvpar", "cx",
                "device", "nvpar_gp", "nvperp_gp", "vpar_sigma_fac", "vperp_sigma_fac", "t_neut", "zns0_neut",
                "ns0_neut")


def _gridgen_in(p):
    g = _abi.KaosGridgenIn()
    for k in GRIDGEN_KEYS:
        if k in p:
            setattr(g, k, p[k])
    if "vperp3d" not in p:
        g.vperp3d = 1
    if "symvpar" not in p:
        g.symvpar = 1
    return g


def gridgen_dims(p):
    """Cell counts of the grids kaos_generate_grid makes from the model inputs
    `p` (the keys of kaos_gridgen_in; host arithmetic, no device)."""
    d = _abi.KaosGridgenDims()
    rc = _abi.load().kaos_gridgen_dims_of(C.byref(_gridgen_in(p)), C.byref(d))
    if rc:
        raise KaosError(rc, "kaos_gridgen_dims_of: invalid grid set-up inputs")
    return {n: getattr(d, n) for n, _ in d._fields_}


def generate_grid(p):
    """Grid set-up of one tube on the device (kaos_generate_grid) from the model
    inputs `p` (ConfigGridGeneratorSub, ConfigGridGeneratorSubB and
    VelGridGeneratorSub, _0_ConfigGridGenerator.f90, _0_1_1_ConfigGridGeneratorB.f90,
    _0_0_VelGridGenerator.f90): dict of the per-cell arrays (GRIDGEN_CFG names
    plus ns_neut), counts (nsnormC0), nsnorm_clb / nsnorm_cub, every axis as
    (GL, GH, GC, d), d3v (2-D) or d3v3 (3-D), ena_hh and ena_d33v [Q][bin],
    and the cell counts of each axis under "ncell"."""
    dims = gridgen_dims(p)
    nq = dims["nq"]
    out = _abi.KaosGridgenOut()
    arrs = {}
    for k in _abi.GRIDGEN_CFG + ("ns_neut",):
        arrs[k] = np.zeros(nq)
        setattr(out, k, ptr(arrs[k]))
    counts = np.zeros(nq, np.int64)
    out.counts = ptr(counts, C.c_int64)
    ncell = dict(zip(("vpar", "vperp", "vperp1", "vperp2", "vp", "vq", "vphi"),
                     (dims["nvpar"], dims["nvperp"], dims["nvperp1"], dims["nvperp2"], dims["nvp"], dims["nvq"],
                      dims["nvphi"])))
    for ax, names in _abi.GRIDGEN_AXES:
        if ncell[ax]:
            for n in names:
                arrs[n] = np.zeros(ncell[ax])
                setattr(out, n, ptr(arrs[n]))
    if dims["nvperp1"]:
        arrs["d3v3"] = np.zeros(dims["nvperp1"] * dims["nvperp2"] * dims["nvpar"])
        out.d3v3 = ptr(arrs["d3v3"])
    else:
        arrs["d3v"] = np.zeros(dims["nvperp"] * dims["nvpar"])
        out.d3v = ptr(arrs["d3v"])
    nb = dims["nvp"] * dims["nvq"] * dims["nvphi"]
    if nb:
        arrs["ena_hh"], arrs["ena_d33v"] = np.zeros(nq * nb), np.zeros(nq * nb)
        out.ena_hh, out.ena_d33v = ptr(arrs["ena_hh"]), ptr(arrs["ena_d33v"])
    rc = _abi.load().kaos_generate_grid(C.byref(_gridgen_in(p)), C.byref(out))
    if rc:
        raise KaosError(rc, "kaos_generate_grid failed")
    res = {k: arrs[k] for k in _abi.GRIDGEN_CFG + ("ns_neut",)}
    res["counts"] = counts
    res["nsnorm_clb"], res["nsnorm_cub"] = out.nsnorm_clb, out.nsnorm_cub
    for ax, names in _abi.GRIDGEN_AXES:
        if ncell[ax]:
            res[ax] = tuple(arrs[n] for n in names)
    for k in ("d3v", "d3v3", "ena_hh", "ena_d33v"):
        if k in arrs:
            res[k] = arrs[k]
    res["ncell"] = ncell
    return res


def read_spinup(directory, species=1, flux_tube=1, nq_cap=1 << 20):
    """kaos_read_spinup: the spin-up files of a previous run as the reference's
    restart reads them (_1_SimParameterization.f90:801-946).  Returns a dict of
    the scalars (nq_lb, nq_ub, nsnorm_clb, nsnorm_cub, lb_replenish,
    ub_replenish, nq) and the arrays density, temperature, ea_inertial,
    ea_pressure, ea_mag."""
    lib = _abi.load()
    st = _abi.KaosSpinupState()
    bufs = [np.zeros(nq_cap) for _ in range(5)]
    rc = lib.kaos_read_spinup(str(directory).encode(), species, flux_tube, nq_cap, C.byref(st),
                              *[ptr(b) for b in bufs])
    if rc:
        raise KaosError(rc, lib.kaos_last_error(None).decode())
    out = {k: getattr(st, k) for k, _ in _abi.KaosSpinupState._fields_}
    for k, b in zip(("density", "temperature", "ea_inertial", "ea_pressure", "ea_mag"), bufs):
        out[k] = b[:st.nq].copy()
    return out
