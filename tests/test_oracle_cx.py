"""Pins of the oracle's charge-exchange restatement (QEXCHANGEflag = 1), on the CPU.

* The collision stream is numpy's Philox4x64-10 with key (seed, "CXNT") and
  counter (cell, master step, tube, block).
* sigmaIonNeutRT (_8_6_6_1:93-114) against numpy's log/exp/pow of the
  reference's expression, its closed form at E = 1 keV and its zero.
* PoissonRNGSub (_3_2_PoissonRNG.f90:30-40): the product-of-uniforms loop
  restated in numpy on numpy's Philox uniforms gives the same counts; mean
  and variance of 4000 draws at three means; mean 0 gives the loop's -1.
* One cell's count (kaos_cx_cell) against a numpy restatement of
  IonNeutralCollisionFrequencyASub: neutral Gaussians (Box-Muller cos
  branch), the ion mean velocity at the cell centre in Cartesian components,
  the relative speed, the Poisson mean, the cap at the cell's macro-ions, the
  -2 codes, and the one-rank share of IonNeutralCollisionFrequencyBSub.
* IonNeutralCollisionsSub's ENA velocity (_8_1_1_1:31-58) against a numpy
  restatement of the reference's lines (Vp = Vperp1 on p-hat, Vphi = Vperp2
  on phi-hat, Vq = b-hat(thetak4) . V); with no perpendicular components and
  thetak4 at the ion's point the ENA keeps the ion's velocity.
* The ballistic ENA push against a numpy RK4 of the same stages, the exact
  straight line without gravity, and energy conservation under gravity.
* An oracle run: each step's conversions are the first nint(poi) ions of
  every cell in id order (raw and filtered moments).
* The ENA velocity components (|Vp|, Vq, Vphi) against the reference's
  trigonometric forms (theta = acos(z/r), phi = atan2(y, x)).
* The ENA distribution function and eight moments against a per-bin numpy
  restatement of _8_4_2 / _8_9 / _8_10_1..8 with non-uniform metric factors
  (1e-12), and closed forms: one occupied bin (M0 = count nsnormfac/d3x, M1 =
  the bin centre, second moments 0), two bins (weighted mean and spread), an
  empty cell (zeros).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from oracle import ptr
from kaos_amd import grids

RE, GM = 6.38e6, 6.67e-11 * 5.98e24
SEED, STEP, TUBE = 0x0123456789ABCDEF, 40, 3


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def _np_uniforms(Q, blk0, nblk, seed=SEED, step=STEP, tube=TUBE):
    """Uniforms of blocks blk0.. of a cell's stream from numpy's Philox."""
    out = []
    for b in range(blk0, blk0 + nblk):
        ctr = np.array([Q - 1, step, tube, b], dtype=np.uint64)     # numpy adds 1 before the block
        g = np.random.Philox(counter=ctr, key=np.array([seed, 0x43584E54], dtype=np.uint64))
        out += [((int(x) >> 11) + 1) / 2.0 ** 53 for x in g.random_raw(4)]
    return out


def test_cx_stream_is_numpy_philox(lib):
    for Q, blk in ((1, 0), (17, 5), (256, 2)):
        u = np.zeros(4)
        lib.orc_cx_block(SEED, STEP, TUBE, Q, blk, ptr(u))
        assert u.tolist() == _np_uniforms(Q, blk, 1)


def _np_sigma(m2):
    if m2 == 0.0:
        return 0.0
    E = (6.242e18 * 1e-3) * m2
    return (4.07 - 0.269 * np.log(E)) ** 2 * (1.0 - np.exp(-415.0 / E)) ** 0.8


def test_cx_sigma(lib):
    for m2 in np.concatenate([np.logspace(-22, -14, 60), [1.0 / (6.242e18 * 1e-3)]]):
        assert abs(lib.orc_cx_sigma(m2) - _np_sigma(m2)) <= 2e-14 * _np_sigma(m2)
    # E = 1 keV: ln E = 0 and exp(-415) is below 2^-53 of 1
    assert abs(lib.orc_cx_sigma(1.0 / (6.242e18 * 1e-3)) - 4.07 ** 2) < 1e-13
    assert lib.orc_cx_sigma(0.0) == 0.0


def _np_poisson(mu, Q):
    L = math.exp(-mu) if mu < 745 else 0.0
    u = _np_uniforms(Q, 2, 1200)
    prod, k, i = 1.0, -1, 0
    while prod > L:
        prod *= u[i]
        i += 1
        k += 1
    return k


def test_poisson_product_loop_and_moments(lib):
    for mu, Q in ((0.3, 1), (4.0, 2), (37.5, 3), (200.0, 4)):
        assert lib.orc_poisson(mu, SEED, STEP, TUBE, Q) == _np_poisson(mu, Q)
    # mean 0 (no neutrals): exp(0) = 1 and the loop never runs, so the
    # reference's variate is -1 (the cell then converts nothing)
    assert all(lib.orc_poisson(0.0, SEED, s, TUBE, 1) == -1 for s in range(20))
    for mu in (0.7, 5.0, 60.0):
        x = np.array([lib.orc_poisson(mu, SEED, s, 0, q) for s in range(40) for q in range(1, 101)])
        se = math.sqrt(mu / len(x))
        assert abs(x.mean() - mu) < 5 * se, mu
        assert abs(x.var() - mu) < 0.15 * mu, mu


def _cell_in(m0=1e9, m1par=250.0, m1p1=30.0, m1p2=-20.0, m2par=2e-20, cth=0.6, nsneut=5e4, nqc=300.0,
             h=0.05, scale=1e-7):
    ell = 1.0 + 3.0 * cth * cth
    sig_n = math.sqrt(1.38e-23 * 848.0 / (16 * 1.66054e-27))
    d3x, nsnormfac = 3.0e15, 1.0e13
    NqRT = nqc * (nsnormfac / d3x)
    return [m0, m1par, m1p1, m1p2, m2par, cth, ell, 0.0, 1.0, nsneut, h, sig_n, NqRT, d3x, nsnormfac, nqc, scale]


def _np_cell(v, Q):
    m0, m1par, m1p1, m1p2, m2par, c, ell, cphi, sphi, nsneut, h, sig_n, NqRT, d3x, nsnormfac, nqc, sc = v
    sigma = sc * _np_sigma(m2par)
    nu = -2.0
    if m0 != 0.0:
        u = _np_uniforms(Q, 0, 2)
        g = [math.sqrt(-2.0 * math.log(u[2 * i])) * math.cos(2.0 * math.pi * u[2 * i + 1]) for i in range(3)]
        Vn = [sig_n * x for x in g]
        s, sl = math.sqrt(1 - c * c), math.sqrt(ell)
        prho = (3.0 * m1par * c * s + m1p1 * (1.0 - 3.0 * c * c)) / sl
        M = [cphi * prho - m1p2 * sphi, sphi * prho + m1p2 * cphi,
             m1par * (3.0 * c * c - 1.0) / sl + 3.0 * m1p1 * c * s / sl]
        vrel = math.sqrt(sum((M[i] - Vn[i]) ** 2 for i in range(3)))
        if sigma != 0.0:
            nu = min(float(_np_poisson(nsneut * sigma * vrel * h, Q)), NqRT * (d3x / nsnormfac))
    poi = min(nu, nqc) if (nu > 0 and nqc != 0) else 0.0
    return sigma, nu, poi


def test_cx_cell_against_restatement(lib):
    cases = [_cell_in(), _cell_in(nsneut=4e5), _cell_in(nqc=7.0), _cell_in(m0=0.0), _cell_in(m2par=0.0),
             _cell_in(cth=-0.8, m1par=-400.0, m1p2=55.0), _cell_in(scale=1.0, nqc=90.0)]
    seen_cap = seen_small = False
    for Q, v in enumerate(cases, start=1):
        out = np.zeros(3)
        lib.orc_cx_cell(ptr(np.array(v)), SEED, STEP, TUBE, Q, ptr(out))
        sg, nu, poi = _np_cell(v, Q)
        assert abs(out[0] - sg) <= 2e-14 * abs(sg), Q
        assert out[1] == nu and out[2] == poi, (Q, out, nu, poi)
        seen_cap |= (nu > 0 and nu == v[12] * (v[13] / v[14]))
        seen_small |= (0 < nu < v[15])
    assert seen_cap and seen_small
    # the -2 codes: an empty cell, and sigma = 0
    for v in (_cell_in(m0=0.0), _cell_in(m2par=0.0)):
        out = np.zeros(3)
        lib.orc_cx_cell(ptr(np.array(v)), SEED, STEP, TUBE, 9, ptr(out))
        assert out[1] == -2.0 and out[2] == 0.0


def _frames(c, phi):
    s, sl = math.sqrt(1 - c * c), math.sqrt(1 + 3 * c * c)
    cp, sp = math.cos(phi), math.sin(phi)
    b = np.array([3 * c * s * cp, 3 * c * s * sp, 3 * c * c - 1]) / sl
    p = np.array([(1 - 3 * c * c) * cp, (1 - 3 * c * c) * sp, 3 * c * s]) / sl
    f = np.array([-sp, cp, 0.0])
    return b, p, f


def test_ena_velocity_from_ion(lib):
    p, phi = 4.0, 1.1
    cphi, sphi = math.cos(phi), math.sin(phi)
    rng = np.random.default_rng(9)
    for rs in (1.3 * RE, -2.1 * RE, 2.9 * RE):
        r = abs(rs)
        cth = math.copysign(math.sqrt(1 - r / (RE * p)), rs)
        for trial in range(4):
            vs, v1, v2 = rng.normal(0, 2e3, 3)
            ck4 = cth + rng.normal(0, 1e-4)
            X = np.zeros(6)
            lib.orc_ena_from_ion(rs, vs, v1, v2, ck4, p, cphi, sphi, 0.0, ptr(X))
            b4, p4, f4 = _frames(ck4, phi)
            bN, _, _ = _frames(cth, phi)
            V = X[3:]
            vq = vs * float(bN @ b4)                 # b-hat(thetak4) . (V|| b-hat(theta_N))
            assert abs(V @ b4 - vq) < 1e-9 and abs(V @ p4 - v1) < 1e-9 and abs(V @ f4 - v2) < 1e-9
            assert abs(V @ V - (vq * vq + v1 * v1 + v2 * v2)) < 1e-8 * (V @ V)
            # position: the ion's Cartesian point
            rho = r * math.sqrt(r / (RE * p))
            assert np.allclose(X[:3], [rho * cphi, rho * sphi, r * cth], rtol=0, atol=1e-6)
        # no perpendicular velocity and the ion's own frame: the ion's velocity
        X = np.zeros(6)
        lib.orc_ena_from_ion(rs, 1234.5, 0.0, 0.0, cth, p, cphi, sphi, 0.0, ptr(X))
        bN, _, _ = _frames(cth, phi)
        assert np.allclose(X[3:], 1234.5 * bN, rtol=0, atol=1e-9)


def _np_ena_push(X, h, grav):
    def acc(x):
        if not grav:
            return np.zeros(3)
        r = math.sqrt(x @ x)
        return -(GM / (r * r)) * (x / r)
    x, V = X[:3].copy(), X[3:].copy()
    a1 = acc(x)
    a2 = acc(x + h / 2 * V)
    a4 = acc(x + h * V + h * h / 2 * a1)
    k2, k4 = V + h / 2 * a1, V + h * a2
    return np.concatenate([x + h / 6 * (V + 2 * k2 + 2 * k2 + k4), V + h / 6 * (a1 + 2 * a2 + 2 * a2 + a4)])


def test_ena_push(lib):
    X0 = np.array([0.3 * RE, 1.7 * RE, -0.9 * RE, 900.0, -1500.0, 2500.0])
    X = X0.copy()
    lib.orc_ena_push(ptr(X), 0.5, 0)                 # no gravity: a straight line
    assert np.allclose(X[:3], X0[:3] + 0.5 * X0[3:], rtol=1e-15, atol=0) and np.array_equal(X[3:], X0[3:])
    X = X0.copy()
    ref = X0.copy()
    e0 = 0.5 * X0[3:] @ X0[3:] - GM / math.sqrt(X0[:3] @ X0[:3])
    for _ in range(200):
        lib.orc_ena_push(ptr(X), 2.0, 1)
        ref = _np_ena_push(ref, 2.0, 1)
    assert np.allclose(X, ref, rtol=1e-12, atol=0)
    e1 = 0.5 * X[3:] @ X[3:] - GM / math.sqrt(X[:3] @ X[:3])
    assert abs(e1 - e0) < 1e-9 * abs(e0)
    # 400 s of fall changed the velocity by ~ g t (a real acceleration, not a no-op)
    assert np.linalg.norm(X[3:] - X0[3:]) > 1000.0


def _cx_tube(scale=1e-7, filt=0, h=0.05):
    t = grids.with_vgrid3(grids.dipole_tube(nq=24, nvpar=20, nvperp=16, n_total=12000, ratio=1.1, h=h), 10, 8,
                          vperp_max=6.0e3)
    return grids.with_charge_exchange(t, ena_cap=20000, sigma_scale=scale, filt=filt)


@pytest.mark.parametrize("filt", [0, 1])
def test_oracle_converts_the_first_ions_of_each_cell(filt):
    t = _cx_tube(filt=filt)
    q, v, w = grids.initial_ions(t, seed=31)
    o = oracle.Oracle(t, q, v, w, ndatfac=2, seed=8, capacity=len(q) + 20 * (t.n_inj_lb + t.n_inj_ub) + 500)
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * np.pi, len(w))
    o.set_vperp(w * np.cos(ph), w * np.sin(ph))
    o.step(1)
    made = 0
    for n in range(2, 11):
        before = o.particles()
        poi = o.cx()["poi"]
        enas_b = set(o.enas()["id"].tolist())
        _, _, idb = o.vperp12()
        o.step(n)
        new = sorted(set(o.enas()["id"].tolist()) - enas_b)
        # the first nint(poi[Q]) ids of each cell before the step
        want = []
        for Q in range(t.nq):
            k = int(np.floor(poi[Q] + 0.5)) if poi[Q] > 0 else 0
            ids_q = np.sort(idb[before["qind"] == Q + 1])
            want += ids_q[:k].tolist()
        assert new == sorted(want), n
        made += len(new)
    assert made > 100
    cx = o.cx()
    assert np.all((cx["nu"] == -2.0) | (cx["nu"] >= 0.0)) and np.all(cx["poi"] >= 0)
    o.close()


# --- ENA velocity space -------------------------------------------------------

def _np_vcomp(X):
    """The reference's ENA components at its point: Vpar = b-hat . V
    (KUA:978-980), Vp = |p-hat . V| and Vphi = phi-hat . V (KUA:1156-1160),
    with theta = acos(z/r) and phi = atan2(y, x)."""
    x, y, z, vx, vy, vz = X
    r = math.sqrt(x * x + y * y + z * z)
    th, ph = math.acos(z / r), math.atan2(y, x)
    c, s, ell = math.cos(th), math.sin(th), 1 + 3 * math.cos(th) ** 2
    vpar = 3 * c * s * (vx * math.cos(ph) + vy * math.sin(ph)) / math.sqrt(ell) + vz * (3 * c * c - 1) / math.sqrt(ell)
    vp = abs((1 - 3 * c * c) * (vx * math.cos(ph) + vy * math.sin(ph)) / math.sqrt(ell) + 3 * vz * c * s / math.sqrt(ell))
    vphi = -vx * math.sin(ph) + vy * math.cos(ph)
    return vp, vpar, vphi


def test_ena_velocity_components(lib):
    rng = np.random.default_rng(21)
    for _ in range(300):
        r = rng.uniform(1.1, 3.0) * RE
        th, ph = rng.uniform(0.2, 2.9), rng.uniform(-3.1, 3.1)
        X = np.array([r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph), r * math.cos(th),
                      *rng.normal(0, 2e3, 3)])
        o3 = np.zeros(3)
        lib.orc_ena_vcomp(ptr(X), ptr(o3))
        assert np.allclose(o3, _np_vcomp(X), rtol=0, atol=1e-9)
        assert o3[0] >= 0.0


def _ena_tube():
    t = _cx_tube()
    return t, oracle.cfg_of(t, 1, 1)


def _np_ena_moments(t, counts, ms):
    """Per-bin restatement of ENADistributionFunctionsSub and the ENA moment
    integrands (_8_4_2:24-36, _8_9:25-35, _8_10_1..8) for every cell."""
    a = t.arrays
    nvp, nvq, nvphi = t.nvp, t.nvq, t.nvphi
    dV = (a["dvp"][:, None, None] * a["dvq"][None, :, None]) * a["dvphi"][None, None, :]
    hh = a["ena_hh"].reshape(t.nq, nvp, nvq, nvphi)
    d33 = a["ena_d33v"].reshape(t.nq, nvp, nvq, nvphi)
    VP, VQ, VF = np.meshgrid(a["vpGC"], a["vqGC"], a["vphiGC"], indexing="ij")
    out = np.zeros((8, t.nq))
    for Q in range(t.nq):
        c = counts[Q]
        if c.sum() == 0:
            continue
        F = (c * (t.nsnormfac / a["d3xC"][Q])) / d33[Q]
        w = dV * hh[Q] * F
        M0 = w.sum()
        m1 = [(w * V).sum() / M0 for V in (VP, VQ, VF)]
        d = [m1[0] - VP, m1[1] - VQ, m1[2] - VF]
        s2 = [(w * (d[0] ** 2 + d[1] ** 2 + d[2] ** 2)).sum()] + [(w * x ** 2).sum() for x in d]
        out[:, Q] = [M0, *m1, *[6.242e18 * (ms / 2) * x / M0 for x in s2]]
    return out


def test_ena_moments_against_restatement(lib):
    t, cfg = _ena_tube()
    # non-uniform metric factors that differ from cell to cell, as the
    # reference's grid generator makes them (they depend on thetaGC)
    rng = np.random.default_rng(4)
    nb = t.nvp * t.nvq * t.nvphi
    hh = rng.uniform(0.5, 2.0, (t.nq, nb))
    t.arrays["ena_hh"] = hh.ravel()
    dV = ((t.arrays["dvp"][:, None, None] * t.arrays["dvq"][None, :, None]) * t.arrays["dvphi"][None, None, :]).ravel()
    t.arrays["ena_d33v"] = (hh * dV[None, :]).ravel()
    cfg = oracle.cfg_of(t, 1, 1)
    counts = np.zeros((t.nq, t.nvp, t.nvq, t.nvphi))
    counts[0] = rng.poisson(3.0, (t.nvp, t.nvq, t.nvphi))           # a populated cell
    counts[3, 2, 5, 4] = 7                                           # one occupied bin
    counts[5, 1, 2, 3], counts[5, 6, 9, 8] = 1, 3                    # two bins
    M = np.zeros(8 * t.nq)
    lib.orc_ena_moments_from_counts(C.byref(cfg), ptr(counts.ravel()), ptr(counts.reshape(t.nq, -1).sum(1)),
                                    ptr(M))
    M = M.reshape(8, t.nq)
    ref = _np_ena_moments(t, counts, t.ms)
    assert np.allclose(M, ref, rtol=1e-12, atol=0)
    a = t.arrays
    # one bin: M0 = count nsnormfac/d3x (hh and d33v cancel), M1 = the bin centre, second moments 0
    assert abs(M[0, 3] - 7 * t.nsnormfac / a["d3xC"][3]) < 1e-12 * M[0, 3]
    assert (M[1, 3], M[2, 3], M[3, 3]) == (a["vpGC"][2], a["vqGC"][5], a["vphiGC"][4])
    assert np.all(M[4:, 3] == 0.0)
    # two bins: the count-weighted mean and spread (equal cell volumes)
    w = np.array([1.0, 3.0])
    vp = np.array([a["vpGC"][1], a["vpGC"][6]])
    m1p = (w * vp).sum() / w.sum()
    assert abs(M[1, 5] - m1p) < 1e-9
    assert abs(M[5, 5] - 6.242e18 * t.ms / 2 * (w * (vp - m1p) ** 2).sum() / w.sum()) < 1e-9 * M[5, 5]
    assert np.all(M[:, 1] == 0.0)                                    # empty cell
