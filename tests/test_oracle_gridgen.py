"""Pins of the grid set-up (ConfigGridGeneratorSub, ConfigGridGeneratorSubB,
VelGridGeneratorSub; oracle.generate_grid, the arithmetic kaos_generate_grid
runs on the device) on the CPU.

Against the reference's own set-up (oracle/_ref/kaos_ref_loop, its Fortran
compiled unmodified, run here with 2 MPI ranks; skipped where it cannot run):
the q edges and centres, the cell geometry and volumes, Ts and Te, the 3-D
velocity grid with d3vCT, and the initial ions per config cell -- the
reference's counts are the truncated hydrostatic counts nsnormC0 (an integer
array) in every cell, and its LB injection is nint of the first cell's count
over the ranks.

Closed forms: the q grid's steps; the root solve's point on the line
(q = RE^2 cos(theta)/r^2, r = L RE sin^2(theta)); the arc length of the
line against the sum of hq dq; the hydrostatic density against the
barometric law in the gravitational potential G ME / r; the symmetric and
one-sided velocity grids; the ENA metric at the equator, where it reduces
to |sin(Vq)|^3 RE, |Vp|^3 / RE^2 and |Vp sin(Vq)|; kaos_sin against numpy.
"""
import math

import numpy as np
import pytest

import oracle
import oracle.ref_loop as rl

RE, GG, ME, KB = 6.38e6, 6.67e-11, 5.98e24, 1.38e-23
MO = 16.0 * 1.66054e-27
NSNORMFAC = 3.85e17

# the harness's model inputs (oracle/ref/loop/kaos_ref_input.f90)
HARNESS = dict(nq_ub=64, qGA=0.7790, qGB=0.0556, L=4.0, phi=math.pi / 2.0, ms=MO, ti=3482.6, te=1e4, ns0=1e8,
               nsnormfac=NSNORMFAC, vperp3d=1, symvpar=1, cx=0, nvpar_gp=14, nvperp_gp=14, vpar_sigma_fac=7.4,
               vperp_sigma_fac=7.4)


@pytest.fixture(scope="module")
def ref(tmp_path_factory):
    if not rl.available():
        if rl.build() is None or not rl.available():
            pytest.skip("reference sources or MPI absent: the reference loop cannot be built or run")
    wd = str(tmp_path_factory.mktemp("kaos_ref_setup"))
    rl.run(wd, 3, NSNORMFAC, ranks=2, seed=5)
    return rl.load_inputs(wd)


def test_config_grid_equals_the_reference_setup(ref):
    t, g = ref["tube"], oracle.generate_grid(HARNESS)
    a = t.arrays
    assert t.nq == len(g["qGL"]) == 64
    for k in ("qGL", "qGH", "qGC", "dqC", "Te"):
        assert np.array_equal(g[k], a[k]), k
    assert np.array_equal(g["Ts"], ref["Ts"])
    # the root solve: the reference's resolvent cubic in complex arithmetic, here its real form
    for k in ("rGC", "ellGC", "hqC", "d3xC"):
        assert np.allclose(g[k], a[k], rtol=1e-13, atol=0), (k, np.max(np.abs(g[k] / a[k] - 1)))
    assert np.allclose(g["cthGC"], np.cos(ref["theta"]), rtol=1e-12, atol=1e-15)


def test_velocity_grid_equals_the_reference_setup(ref):
    t, g = ref["tube"], oracle.generate_grid(HARNESS)
    a = t.arrays
    for ax, pre in (("vperp1", "v1"), ("vperp2", "v2")):
        for k, suf in zip(g[ax], ("GL", "GH", "GC", "d")):
            name = ("d" + pre) if suf == "d" else pre + suf
            assert np.array_equal(k, a[name]), name
    for k, name in zip(g["vpar"], ("vparGL", "vparGH", "vparGC", "dvpar")):
        assert np.array_equal(k, a[name]), name
    assert np.array_equal(g["d3v3"], a["d3v3"].ravel())


def test_initial_counts_are_the_truncated_hydrostatic_counts(ref):
    t, g = ref["tube"], oracle.generate_grid(HARNESS)
    qGC = t.arrays["qGC"]
    Q = np.array([int(np.argmin(np.abs(qGC - q))) for q in ref["ions"]["q0"]])
    got = np.bincount(Q, minlength=t.nq)
    assert np.array_equal(got, g["counts"])
    x = (g["density"] / NSNORMFAC) * g["d3xC"]
    assert np.any(x - np.floor(x) > 0.5)                      # rounding to nearest would differ there
    assert ref["n_inj_rank"] == int(math.floor(g["counts"][0] / ref["ranks"] + 0.5))


def _dipole_arc(L, c):
    """Arc length of a dipole line from the equator to cos(theta) = c (closed form)."""
    s3 = math.sqrt(3.0)
    return L * RE * (s3 * c * math.sqrt(1.0 + 3.0 * c * c) + math.asinh(s3 * c)) / (2.0 * s3)


@pytest.mark.parametrize("qGA,qGB", [(0.7790, 0.0556), (-0.7790, -0.0556)])
def test_config_grid_closed_forms(qGA, qGB):
    p = dict(HARNESS, qGA=qGA, qGB=qGB, nq_ub=96)
    g = oracle.generate_grid(p)
    lib = oracle.load()
    ngp = 96 + 3
    pts = np.array([lib.orc_gg_qpoint(qGA, qGB, ngp, i) for i in range(1, ngp + 1)])
    step = abs(qGB - qGA) / ngp
    assert np.allclose(np.abs(np.diff(pts[:-1])), step, rtol=1e-12) and pts[-1] == qGB and pts[0] == qGA
    assert np.array_equal(g["qGL"], pts[1:97]) and np.array_equal(g["qGH"], pts[2:98])
    r, c, L = g["rGC"], g["cthGC"], p["L"]
    s2 = 1.0 - c * c
    assert np.allclose(RE * RE * c / (r * r), g["qGC"], rtol=1e-13)
    assert np.allclose(r, L * RE * s2, rtol=1e-13)
    assert np.allclose(g["ellGC"], 1.0 + 3.0 * c * c, rtol=1e-15)
    # sum of hq dq over the cells against the arc length between the outer edges
    def cos_of(q):
        rr = np.array([oracle.load().orc_root_r(x, L) for x in q])
        cc = np.sqrt(1.0 - rr / (L * RE))
        return np.where(np.asarray(q) <= 0, -cc, cc)
    ce = cos_of([g["qGL"][0], g["qGH"][-1]])
    arc = abs(_dipole_arc(L, ce[0]) - _dipole_arc(L, ce[1]))
    assert abs(np.sum(g["hqC"] * g["dqC"]) / arc - 1.0) < 1e-3              # midpoint rule
    # volumes: hq hp hphi dq dp dphi with hp = RE sin^3 / sqrt(ell), hphi = r sin
    sn = np.sqrt(s2)
    d3x = g["hqC"] * (RE * sn ** 3 / np.sqrt(g["ellGC"])) * (r * sn) * g["dqC"] * 1e-3 * (math.pi / 180.0)
    assert np.allclose(g["d3xC"], d3x, rtol=1e-13)


def test_hydrostatic_density_is_barometric():
    """sum of ICbbp = the field-aligned gravity integrated along the line: the
    density follows exp(-ms (Phi(r) - Phi(r_1)) / (kB (Ti + Te))), Phi = -G ME / r,
    from the upper edge of the ghost cell, to the midpoint rule's accuracy."""
    p = dict(HARNESS, nq_ub=128)
    g = oracle.generate_grid(p)
    lib = oracle.load()
    q0 = lib.orc_gg_qpoint(p["qGA"], p["qGB"], 128 + 3, 2)          # qGH of set-up cell 1
    r0 = lib.orc_root_r(q0, p["L"])
    rh = np.array([lib.orc_root_r(q, p["L"]) for q in g["qGH"]])
    want = -(p["ms"] * GG * ME * (1.0 / r0 - 1.0 / rh)) / (KB * (p["ti"] + p["te"]))
    got = np.log(g["density"] / p["ns0"])
    assert np.max(np.abs(got - want)) < 1e-4 * np.max(np.abs(want))
    assert np.array_equal(g["counts"], np.trunc((g["density"] / p["nsnormfac"]) * g["d3xC"]).astype(np.int64))


def test_neutral_density_and_counts():
    p = dict(HARNESS, cx=1, t_neut=848.0, zns0_neut=RE + 719e3, ns0_neut=1.423e11, nsnormfac=1e13)
    g = oracle.generate_grid(p)
    r = g["rGC"]
    n = p["ns0_neut"] * np.exp(-p["ms"] * (GG * ME / (KB * p["t_neut"])) * (1.0 / p["zns0_neut"] - 1.0 / r))
    x = (n / p["nsnormfac"]) * g["d3xC"]
    assert np.all(np.abs(g["ns_neut"] - np.trunc(x)) <= 1.0) and np.all(g["ns_neut"] == np.trunc(g["ns_neut"]))
    assert g["ns_neut"][0] > 0


@pytest.mark.parametrize("sym", [1, 0])
@pytest.mark.parametrize("south", [0, 1])
def test_ion_velocity_grids(sym, south):
    p = dict(HARNESS, symvpar=sym, vperp3d=0, nvpar_gp=20, nvperp_gp=16)
    if south:
        p.update(qGA=-0.7790, qGB=-0.0556)
    g = oracle.generate_grid(p)
    sig = math.sqrt(KB * p["ti"] / p["ms"])
    GB = p["vpar_sigma_fac"] * sig
    gl, gh, gc, d = g["vpar"]
    assert len(gl) == 18 and np.array_equal(gl[1:], gh[:-1]) and np.all(d > 0)
    if sym:
        N = 10
        assert gl[0] == -GB and gh[-1] == GB and gh[N - 2] == 0.0             # -GB .. 0 .. GB
        assert np.allclose(d[: N - 2], GB / N, rtol=1e-12) and abs(d[N - 2] / (2 * GB / N) - 1) < 1e-12
        assert np.array_equal(gl, -gh[::-1])
    else:
        N = 20
        p1 = np.array([-4.0 * sig + (i - 1.0) * ((GB + 4.0 * sig) / N) for i in range(1, N)] + [GB])
        edges = -p1[::-1][: N - 1] if south else p1[: N - 1]        # N - 1 of the N points
        assert np.array_equal(np.append(gl, gh[-1]), edges)
    assert np.all(gc != 0.0) and np.array_equal(gc[gc != 1e-15], ((gl + gh) / 2)[gc != 1e-15])
    # the 2-D Vperp grid: symmetric, d3v = 2 pi |Vperp_c| dVperp dV||
    vl, vh, vc, dv = g["vperp"]
    assert vl[0] == -p["vperp_sigma_fac"] * sig and len(vl) == 14
    d3 = (((np.abs(vc) * dv)[:, None] * d[None, :]) * (2.0 * math.pi)).ravel()
    assert np.allclose(g["d3v"], d3, rtol=1e-15)


def test_ena_grid_and_metric_at_the_equator():
    p = dict(HARNESS, cx=1, t_neut=848.0, zns0_neut=RE + 719e3, ns0_neut=1.423e11, nvpar_gp=12, nvperp_gp=10)
    g = oracle.generate_grid(p)
    sn = math.sqrt(KB * 848.0 / p["ms"])
    GB = p["vperp_sigma_fac"] * sn
    N = 5
    p1 = [0.0 + (i - 1.0) * (GB / N) if i != N else GB for i in range(1, N + 1)]
    pts = [-p1[N - i] for i in range(1, N + 1)] + p1[: N - 1]      # -GB .. -0, then P1(i - N)
    vl, vh, vc, dv = g["vp"]
    assert np.array_equal(np.append(vl, vh[-1]), pts) and len(vl) == 8
    assert dv[N - 1] == 0.0                                       # the zero-width cell (-0, 0)
    nb = 8 * 10 * 8
    assert g["ena_hh"].shape == (64 * nb,) and np.all(g["ena_d33v"] > 0)
    lib = oracle.load()
    o2 = np.zeros(2)
    rng = np.random.default_rng(2)
    for vp, vq in rng.uniform(-3e4, 3e4, (50, 2)):
        lib.orc_gg_ena_metric(0.0, 1.0, 1.0, vp, vq, 2.0, 3.0, 5.0, oracle.ptr(o2))
        hh = (RE * abs(math.sin(vq)) ** 3) * (abs(vp) ** 3 / RE ** 2) * abs(vp * math.sin(vq))
        assert abs(o2[0] / hh - 1.0) < 1e-9 and abs(o2[1] / (hh * 30.0) - 1.0) < 1e-9
    # loop cell Q's metric is the one of set-up cell Q's theta (one below its own centre)
    cth1 = g["cthGC"][0]
    b = 3 * 10 * 8 + 4 * 8 + 2
    lib.orc_gg_ena_metric(cth1, math.sqrt(1 - cth1 ** 2), 1 + 3 * cth1 ** 2, g["vp"][2][3], g["vq"][2][4],
                          g["vp"][3][3], g["vq"][3][4], g["vphi"][3][2], oracle.ptr(o2))
    assert g["ena_hh"][nb + b] == o2[0]


def test_sin_against_numpy():
    lib = oracle.load()
    x = np.concatenate([np.linspace(-10, 10, 2001), np.random.default_rng(1).uniform(-2e5, 2e5, 4000)])
    got = np.array([lib.orc_sin(v) for v in x])
    assert np.max(np.abs(got - np.sin(x))) < 4e-16
