"""Flux-tube grids and initial ions for the benchmark and test workloads.

The reference builds its grids in SimParameterizationSub (uniform in q,
_0_ConfigGridGenerator.f90:299-310; velocity grids _0_0_VelGridGenerator.f90)
and places every initial ion at a cell centre (_4_DensityProfileB.f90:40-46).
The workloads here are specified differently (arc-length geometric cells,
exponential density, uniform velocity grids), so they are generated on the
host with numpy and passed to the hot path as inputs, exactly as the Fortran
driver would pass its own arrays.

Dipole field line of L-shell p: r = p RE sin^2(theta), q = RE^2 cos(theta)/r^2.
Arc length from the lower boundary: s = p RE [F(u_LB) - F(u)], u = cos(theta),
F(u) = u sqrt(1+3u^2)/2 + asinh(sqrt(3) u)/(2 sqrt(3)).
"""
from dataclasses import dataclass, field
import math
import numpy as np

RE = 6.38e6
KB = 1.38e-23
MO = 16.0 * 1.66054e-27      # O+ mass, _ModelInput.f90:186
MH = 1.66054e-27             # H+ mass (one atomic mass unit, as the O+ mass is written)
QO = 1.602e-19               # O+ charge, _ModelInput.f90:187
MDIP = 8.06e15
EV_K = 1.602e-19 / KB        # kelvin per eV with the reference's constants


def _F(u):
    return u * np.sqrt(1.0 + 3.0 * u * u) / 2.0 + np.arcsinh(math.sqrt(3.0) * u) / (2.0 * math.sqrt(3.0))


def _u_of_s(s, p, u_lb):
    """cos(theta) at arc length s above the lower boundary (Newton on F)."""
    target = _F(u_lb) - s / (p * RE)
    u = np.full_like(np.asarray(s, dtype=float), u_lb) - s / (p * RE * 2.0)
    for _ in range(60):
        f = _F(u) - target
        du = f / np.sqrt(1.0 + 3.0 * u * u)
        u = u - du
        if np.max(np.abs(du)) < 1e-16:
            break
    return u


@dataclass
class Tube:
    """One flux tube: the kaos_grid fields plus host-side helpers."""
    nq: int
    nvpar: int
    nvperp: int
    p: float
    phi: float
    h: float
    ms: float
    qs: float
    nsnormfac: float
    Ts_lb: float
    Ts_ub: float
    n_inj_lb: int
    n_inj_ub: int
    arrays: dict = field(default_factory=dict)
    s_centre: np.ndarray = None
    ds: np.ndarray = None
    grav: int = 1
    eamb: int = 1
    eambsign: int = 1
    filter_pt: int = 11
    filter_pt_m1par: int = 0          # 0: the same window as filter_pt
    lbcond: int = 1
    ubcond: int = 1
    # ICR wave heating (stochastic); defaults are the reference's
    # _ModelInput.f90:214-229 values
    icr: int = 0
    icr_lambda_perp: float = 0.0
    icr_eta_lh: float = 0.125
    icr_xi_perp1: float = 0.5
    icr_xi_perp2: float = 0.5
    icr_s0: float = 5e-5
    icr_f0: float = 6.5
    icr_chi_perp1: float = 2.1
    icr_chi_perp2: float = 2.1
    # 3-D velocity grid (ION2VPERPflag = 1), set by with_vgrid3
    v3d: int = 0
    nv1: int = 0
    nv2: int = 0
    # charge exchange (QEXCHANGEflag = 1), set by with_charge_exchange
    cx: int = 0
    cx_filter: int = 0
    cx_filter_pt: int = 11
    t_neut: float = 848.0
    m_neut: float = 0.0
    cx_sigma_scale: float = 1.0
    ena_cap: int = 0
    nvp: int = 0
    nvq: int = 0
    nvphi: int = 0

    @property
    def cphi(self):
        return math.cos(self.phi)

    @property
    def sphi(self):
        return math.sin(self.phi)


def velocity_grid(nvpar, nvperp, vpar_max=1.0e4, vperp_max=1.0e4):
    """Uniform V|| in [-vpar_max, vpar_max] and Vperp in [0, vperp_max] with
    contiguous edges; d3v = 2 pi |Vperp_c| dVperp dVpar (_0_0_VelGridGenerator.f90:460-464)."""
    e = -vpar_max + (2.0 * vpar_max) * np.arange(nvpar + 1) / nvpar
    w = vperp_max * np.arange(nvperp + 1) / nvperp
    vparGL, vparGH = e[:-1].copy(), e[1:].copy()
    vperpGL, vperpGH = w[:-1].copy(), w[1:].copy()
    vparGC = (vparGL + vparGH) / 2.0
    vparGC[vparGC == 0.0] = 1e-15          # _0_0_VelGridGenerator.f90:1000
    vperpGC = (vperpGL + vperpGH) / 2.0
    dvpar, dvperp = vparGH - vparGL, vperpGH - vperpGL
    d3v = ((np.abs(vperpGC)[:, None] * dvperp[:, None]) * dvpar[None, :]) * (2.0 * math.pi)
    return dict(vparGL=vparGL, vparGH=vparGH, vparGC=vparGC, dvpar=dvpar, vperpGL=vperpGL,
                vperpGH=vperpGH, vperpGC=vperpGC, dvperp=dvperp, d3v=np.ascontiguousarray(d3v))


def tube_from_setup(g, p, h, qs, ranks=1, ubcond=0, ena_cap=0, sigma_scale=1.0):
    """A Tube on the grids of the reference's set-up: `g` the dict of
    solver.generate_grid (kaos_generate_grid, on the device) for the model
    inputs `p` (the keys of kaos_gridgen_in), time step h [s], charge qs [C].
    Initial ions per cell: g["counts"] (cell_counts); LB injection per master
    step: nint(nsnormCLBG / ranks) (_8_0_7_BoundaryConditions.f90:747-750), UB
    likewise when ubcond.  On the 3-D grid the tube also carries the 2-D arrays
    kaos_grid takes, a uniform Vperp grid over |Vperp1| (counted only in the
    2-D form).  cx: the neutral counts and the ENA grid with its metric;
    ena_cap ENA slots, sigma_scale the cross-section factor."""
    nq = len(g["qGL"])
    nint = lambda x: int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)
    a = {k: np.ascontiguousarray(g[k]) for k in ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC",
                                                    "Te")}
    for n, k in zip(("vparGL", "vparGH", "vparGC", "dvpar"), g["vpar"]):
        a[n] = k
    v3 = "vperp1" in g
    if v3:
        for pre, ax in (("v1", "vperp1"), ("v2", "vperp2")):
            for suf, k in zip(("GL", "GH", "GC"), g[ax][:3]):
                a[pre + suf] = k
            a["d" + pre] = g[ax][3]
        a["d3v3"] = np.ascontiguousarray(g["d3v3"])
        nv1 = len(g["vperp1"][0])
        vmax = float(max(abs(g["vperp1"][0][0]), abs(g["vperp1"][1][-1])))
        vg = velocity_grid(len(a["vparGL"]), nv1, vpar_max=1.0, vperp_max=vmax)
        for n in ("vperpGL", "vperpGH", "vperpGC", "dvperp"):
            a[n] = vg[n]
        a["d3v"] = ((np.abs(vg["vperpGC"])[:, None] * vg["dvperp"][:, None]) * a["dvpar"][None, :]) * (2.0 * math.pi)
        nvperp = nv1
    else:
        for n, k in zip(("vperpGL", "vperpGH", "vperpGC", "dvperp"), g["vperp"]):
            a[n] = k
        a["d3v"] = np.ascontiguousarray(g["d3v"])
        nvperp = len(g["vperp"][0])
    t = Tube(nq=nq, nvpar=len(a["vparGL"]), nvperp=nvperp, p=float(p["L"]), phi=float(p.get("phi", math.pi / 2)),
             h=float(h), ms=float(p["ms"]), qs=float(qs), nsnormfac=float(p["nsnormfac"]), Ts_lb=float(g["Ts"][0]),
             Ts_ub=float(g["Ts"][-1]), n_inj_lb=nint(g["nsnorm_clb"] / ranks),
             n_inj_ub=nint(g["nsnorm_cub"] / ranks) if ubcond else 0,
             arrays={k: np.ascontiguousarray(v, dtype=np.float64) for k, v in a.items()})
    t.ubcond = int(ubcond)
    t.cell_counts = np.asarray(g["counts"], dtype=np.int64)
    if v3:
        t.v3d, t.nv1, t.nv2 = 1, len(g["vperp1"][0]), len(g["vperp2"][0])
    if "vp" in g:
        t.cx, t.t_neut, t.m_neut, t.cx_sigma_scale = 1, float(p["t_neut"]), float(p["ms"]), float(sigma_scale)
        t.arrays["nsneut"] = np.ascontiguousarray(g["ns_neut"])
        for ax in ("vp", "vq", "vphi"):
            for suf, k in zip(("GL", "GH", "GC"), g[ax][:3]):
                t.arrays[ax + suf] = np.ascontiguousarray(k)
            t.arrays["d" + ax] = np.ascontiguousarray(g[ax][3])
        t.nvp, t.nvq, t.nvphi = len(g["vp"][0]), len(g["vq"][0]), len(g["vphi"][0])
        t.arrays["ena_hh"] = np.ascontiguousarray(g["ena_hh"])
        t.arrays["ena_d33v"] = np.ascontiguousarray(g["ena_d33v"])
        t.ena_cap = int(ena_cap)
    return t


def with_vgrid3(tube, nv1, nv2, vperp_max=None):
    """Give a tube the 3-D velocity grid of ION2VPERPflag = 1: Vperp1 and
    Vperp2 uniform in [-vperp_max, vperp_max] (default: the tube's Vpar range)
    with contiguous edges, the tube's Vpar cells, and d3v = (dV1 dV2) dVpar
    as [V1][V2][Vpar] (_0_0_VelGridGenerator.f90:384-398).  Returns the tube."""
    a = tube.arrays
    vmax = float(a["vparGH"][-1]) if vperp_max is None else vperp_max
    for name, n in (("v1", nv1), ("v2", nv2)):
        e = -vmax + (2.0 * vmax) * np.arange(n + 1) / n
        a[name + "GL"], a[name + "GH"] = e[:-1].copy(), e[1:].copy()
        a[name + "GC"] = (a[name + "GL"] + a[name + "GH"]) / 2.0
        a["d" + name] = a[name + "GH"] - a[name + "GL"]
    d3 = ((a["dv1"][:, None, None] * a["dv2"][None, :, None]) * a["dvpar"][None, None, :])
    a["d3v3"] = np.ascontiguousarray(d3)
    tube.v3d, tube.nv1, tube.nv2 = 1, nv1, nv2
    return tube


def with_charge_exchange(tube, n0=1.423e11, r0=None, H=2.0e5, t_neut=848.0, filt=0, filt_pt=11, ena_cap=None,
                         sigma_scale=1.0, ena_grid=(8, 12, 10)):
    """Switch on ion-neutral charge exchange (QEXCHANGEflag = 1; the tube needs
    the 3-D grid).  The neutrals (mass of the tube's species, TNeut kelvin)
    have an exponential density n0 exp(-(r - r0)/H) at the cell centres (H =
    200 km; n0 and r0 = RE + 719 km the reference's ns0NeutIC and zns0NeutIC,
    _ModelInput.f90:211-212); nsnormCNeutGT = n / nsnormfac d3x per cell
    (_0_1_1_ConfigGridGeneratorB.f90:113-114).  filt: MOMENTFILTERflag with
    the M1Perp1/M1Perp2/M2Par window filt_pt.  ena_cap: ENA slots (default:
    the initial ions).  sigma_scale multiplies the cross-section (1: the
    reference's expression as written).  ena_grid: (Vp, Vq, Vphi) cells of
    the ENA velocity grid.  Returns the tube."""
    if not tube.v3d:
        raise ValueError("charge exchange needs the 3-D velocity grid (with_vgrid3)")
    r0 = RE + 7.19e5 if r0 is None else r0
    a = tube.arrays
    n = n0 * np.exp(-(a["rGC"] - r0) / H)
    a["nsneut"] = (n / tube.nsnormfac) * a["d3xC"]
    tube.cx, tube.cx_filter, tube.cx_filter_pt = 1, int(filt), int(filt_pt)
    tube.t_neut, tube.m_neut = float(t_neut), float(tube.ms)
    tube.cx_sigma_scale = float(sigma_scale)
    # ENA velocity grid: |Vp| in [0, vmax], Vq and Vphi in [-vmax, vmax], the
    # Vpar range of the tube; unit metric factors (d33v = dVp dVq dVphi)
    vmax = float(a["vparGH"][-1])
    for name, n, lo in (("vp", ena_grid[0], 0.0), ("vq", ena_grid[1], -vmax), ("vphi", ena_grid[2], -vmax)):
        e = lo + (vmax - lo) * np.arange(n + 1) / n
        a[name + "GL"], a[name + "GH"] = e[:-1].copy(), e[1:].copy()
        a[name + "GC"] = (a[name + "GL"] + a[name + "GH"]) / 2.0
        a["d" + name] = a[name + "GH"] - a[name + "GL"]
    tube.nvp, tube.nvq, tube.nvphi = ena_grid
    # per config cell and bin [Q][ip][iq][iphi]
    dV = ((a["dvp"][:, None, None] * a["dvq"][None, :, None]) * a["dvphi"][None, None, :]).ravel()
    a["ena_hh"] = np.ones(tube.nq * dV.size)
    a["ena_d33v"] = np.tile(dV, tube.nq)
    tube.ena_cap = int(tube.cell_counts.sum()) if ena_cap is None else int(ena_cap)
    return tube


def dipole_tube(nq=64, p=4.0, r_lb=1.05, r_ub=3.0, ratio=1.05, south=False, nvpar=40, nvperp=40,
                phi=math.pi / 2.0, h=0.05, T_eV=0.3, Te=1.0e4, n_total=100000, H=5.0e5,
                ndatfac=1, n_lb_density=1e10, area0=1e6, ms=MO, vmax=1.0e4):
    """Tube from r_lb to r_ub (RE) on L-shell p with nq cells whose arc lengths grow
    geometrically by `ratio`; initial density profile ~ exp(-s/H)."""
    u_lb = math.sqrt(1.0 - r_lb / p)
    u_ub = math.sqrt(1.0 - r_ub / p)
    S = p * RE * (_F(u_lb) - _F(u_ub))
    k = np.arange(nq + 1, dtype=float)
    w0 = S * (ratio - 1.0) / (ratio ** nq - 1.0) if ratio != 1.0 else S / nq
    s_edge = w0 * (ratio ** k - 1.0) / (ratio - 1.0) if ratio != 1.0 else w0 * k
    s_edge[-1] = S
    s_cen = 0.5 * (s_edge[:-1] + s_edge[1:])
    u_e = _u_of_s(s_edge, p, u_lb)
    u_c = _u_of_s(s_cen, p, u_lb)
    sign = -1.0 if south else 1.0
    u_e, u_c = sign * u_e, sign * u_c
    r_e = p * RE * (1.0 - u_e ** 2)
    r_c = p * RE * (1.0 - u_c ** 2)
    q_e = RE ** 2 * u_e / r_e ** 2
    q_c = RE ** 2 * u_c / r_c ** 2
    ell_c = 1.0 + 3.0 * u_c ** 2
    hq = r_c ** 3 / (RE ** 2 * np.sqrt(ell_c))
    ds = np.diff(s_edge)
    B_c = MDIP * np.sqrt(ell_c) / r_c ** 3
    d3x = area0 * (B_c[0] / B_c) * ds
    qGL, qGH = q_e[:-1].copy(), q_e[1:].copy()
    dq = np.abs(qGL - qGH)
    vg = velocity_grid(nvpar, nvperp, vmax, vmax)
    Ts = T_eV * EV_K
    sigma = math.sqrt(KB * Ts / ms)
    # macro-ions per cell ~ exp(-s/H) d3x, normalised to n_total
    wgt = np.exp(-s_cen / H) * d3x
    nc = np.floor(n_total * wgt / wgt.sum()).astype(np.int64)
    nc[0] += n_total - nc.sum()
    nsnormfac = n_lb_density * d3x[0] / max(nc[0], 1)
    dt_m = ndatfac * h
    flux_frac = sigma / math.sqrt(2.0 * math.pi) * dt_m
    n_inj_lb = int(round(nc[0] * flux_frac / ds[0]))
    n_inj_ub = int(round(nc[-1] * flux_frac / ds[-1]))
    arrays = dict(qGL=qGL, qGH=qGH, qGC=q_c, hqC=hq, dqC=dq, d3xC=d3x, rGC=r_c, cthGC=u_c,
                  ellGC=ell_c, Te=np.full(nq, Te))
    arrays.update(vg)
    t = Tube(nq=nq, nvpar=nvpar, nvperp=nvperp, p=p, phi=phi, h=h, ms=ms, qs=QO,
             nsnormfac=nsnormfac, Ts_lb=Ts, Ts_ub=Ts, n_inj_lb=n_inj_lb, n_inj_ub=n_inj_ub,
             arrays={k2: np.ascontiguousarray(v, dtype=np.float64) for k2, v in arrays.items()},
             s_centre=s_cen, ds=ds)
    t.cell_counts = nc
    t.sigma = sigma
    return t


def initial_ions(tube, seed=0, counts=None):
    """Initial (q0, Vpar0, Vperp0): q uniform inside each cell, a Maxwellian of
    the tube's ion temperature (Vperp = |(Vperp1, Vperp2)|)."""
    rng = np.random.default_rng(seed)
    nc = tube.cell_counts if counts is None else counts
    n = int(nc.sum())
    a = tube.arrays
    cell = np.repeat(np.arange(tube.nq), nc)
    f = rng.random(n)
    q0 = a["qGL"][cell] + f * (a["qGH"][cell] - a["qGL"][cell])
    sig = math.sqrt(KB * tube.Ts_lb / tube.ms)
    vpar0 = sig * rng.standard_normal(n)
    v1 = sig * rng.standard_normal(n)
    v2 = sig * rng.standard_normal(n)
    vperp0 = np.sqrt(v1 * v1 + v2 * v2)
    return q0, vpar0, vperp0


# --- the workloads -----------------------------------------------------------

def config1(n_total=100_000):
    """CPU-reference case: L=4, 1.05-3 RE, 64 cells (ratio 1.05), 40x40, 1e5 O+, T=0.3 eV,
    dt=0.05 s, gravity + ambipolar, LB and UB injection."""
    return dipole_tube(nq=64, nvpar=40, nvperp=40, n_total=n_total, ratio=1.05)


def config2(n_total=100_000_000):
    """Full-size single tube: 256 cells (same stretching profile as config 1: ratio
    1.05^(64/256) per cell), 64x64 velocity grid, 1e8 O+."""
    return dipole_tube(nq=256, nvpar=64, nvperp=64, n_total=n_total, ratio=1.05 ** (64.0 / 256.0))


def config4_tubes(ntubes=512, seed=4, scale=1.0):
    """Batched case: L ~ U[3,8], nq ~ U{64..256}, ions log-U[1e4,1e6] (x scale), T ~ U[0.1,1] eV.
    The tube spans 1.05 RE to min(3, 0.8 L) RE."""
    rng = np.random.default_rng(seed)
    tubes = []
    for _ in range(ntubes):
        L = rng.uniform(3.0, 8.0)
        nq = int(rng.integers(64, 257))
        n = int(round(10 ** rng.uniform(4.0, 6.0) * scale))
        T = rng.uniform(0.1, 1.0)
        tubes.append(dipole_tube(nq=nq, p=L, r_ub=min(3.0, 0.8 * L), nvpar=64, nvperp=64,
                                 n_total=max(n, 2 * nq), T_eV=T, ratio=1.05 ** (64.0 / nq)))
    return tubes


def config5_tubes(n_h=50_000_000, n_o=150_000_000, nv1=32, nv2=32, cx=True, sigma_scale=2e-10,
                  ena_grid=(16, 64, 16)):
    """Full physics: one L=6 tube (1.05-3 RE, 256 cells) holding H+ (n_h ions,
    0.5 eV, V grid +-40 km/s) and O+ (n_o ions, 0.3 eV, +-10 km/s) as two
    tubes of one context, dt = 0.02 s; stochastic ICR heating with the
    reference's parameters except spectral indices 2 (diffusion coefficient
    ~ B^-2); the 3-D velocity grid nv1 x nv2 x 64 (Vperp1, Vperp2, Vpar: 65536
    cells per config cell, the count of a 64 x 64 x 16 (Vpar, Vperp, phi)
    grid; nv1 = 0: the 2-D 64 x 64 grid); charge exchange with neutrals of
    the species' mass, 848 K and an exponential density of 200 km scale
    height, the cross-section scaled by sigma_scale, ENAs on a (|Vp|, Vq,
    Vphi) grid of ena_grid cells with slots for 8% of the ions."""
    tubes = []
    for ms, n, T, vmax in ((MH, n_h, 0.5, 4.0e4), (MO, n_o, 0.3, 1.0e4)):
        t = dipole_tube(nq=256, p=6.0, r_ub=3.0, nvpar=64, nvperp=64, n_total=n, T_eV=T, h=0.02,
                        ratio=1.05 ** (64.0 / 256.0), ms=ms, vmax=vmax)
        t.icr = 1
        t.icr_chi_perp1 = t.icr_chi_perp2 = 2.0
        if nv1:
            with_vgrid3(t, nv1, nv2, vperp_max=vmax)
            if cx:
                with_charge_exchange(t, H=2.0e5, sigma_scale=sigma_scale, ena_grid=ena_grid,
                                     ena_cap=max(1024, int(0.08 * n)))
        tubes.append(t)
    return tubes
