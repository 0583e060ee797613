"""Ion-neutral charge exchange (QEXCHANGEflag = 1) on the device through the
C-ABI, against the oracle.

FP64 bit-exact after every step: every ion's state and indices, the ENAs
(Cartesian state, id, Qindk1, in id order), and at master steps the 3-D
counts and moments, the collision counts (sigmaIonNeutRT, nuIonNeutRT,
nuIonNeutPoiT), the ENA counts per cell, the ENA phase-space counts on the
(|Vp|, Vq, Vphi) grid with the eight ENA moments, and the LB/UB ENA escapes
and fluxes:
- a 3-D tube, ndatfac 2, raw moments (MOMENTFILTERflag = 0) and filtered ones
  (= 1, M1Par window 9, M1Perp1/M1Perp2/M2Par window 7), with heating-free
  pushes and the seeded injection stream;
- a long time step (2 s, 30 steps) in which ENAs fall out of the lower
  boundary, with a reorder every master step (cos(thetak4) moves with the
  ions);
- 40000 ions, whose ids need two digits of the device's id select.
Mixed form: the same tube; its moments differ from FP64's by ~1e-7, so a
Poisson count may differ by one where the running product sits at its
threshold: totals of conversions within 2% of the oracle's, ENA counts per
cell within 3 + 2%.  Refusals: charge exchange without the 3-D grid, the
handover and restart files of a charge-exchange context.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
ENA = ("x", "y", "z", "Vx", "Vy", "Vz", "id", "qind")
MASTER = ("nq_counts", "NqRT", "M0", "M1par", "M2", "M2par", "EA", "EG", "esc")


def _tube(h=0.05, filt=0, nq=24, ratio=1.1, n_total=12000):
    t = grids.with_vgrid3(grids.dipole_tube(nq=nq, nvpar=20, nvperp=16, n_total=n_total, ratio=ratio, h=h), 10, 8,
                          vperp_max=6.0e3)
    if filt:
        t.filter_pt_m1par = 9
    return grids.with_charge_exchange(t, ena_cap=max(20000, n_total), sigma_scale=1e-7, filt=filt, filt_pt=7)


def _pair(t, precision, ndatfac, nsteps, reorder=0):
    q, v, w = grids.initial_ions(t, seed=31)
    cap = len(q) + (nsteps + 4) * (t.n_inj_lb + t.n_inj_ub) + 512
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * np.pi, len(w))
    o = oracle.Oracle(t, q, v, w, ndatfac=ndatfac, seed=8, capacity=cap)
    o.set_vperp(w * np.cos(ph), w * np.sin(ph))
    ks = KineticSolver([t], precision=precision, ndatfac=ndatfac, seed=8, capacity=[cap], reorder_every=reorder)
    ks.upload(0, q, v, w)
    ks.upload_vperp(0, w * np.cos(ph), w * np.sin(ph))
    return o, ks


def _same(a, b, keys, what):
    for k in keys:
        assert len(a[k]) == len(b[k]), (what, k, len(a[k]), len(b[k]))
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (what, k)


def _check_step(o, ks, n, ndatfac):
    _same(o.particles(), ks.particles(0), PART, f"ions step {n}")
    _same(o.enas(), ks.enas(0), ENA, f"ENAs step {n}")
    if n == 1 or n % ndatfac == 0:
        _same(o.master(), ks.master(0), MASTER, f"master {n}")
        a3, b3 = o.master3(), ks.master3(0)
        _same(a3, b3, ("n3_counts", "M0", "M1perp1", "M1perp2", "M1par", "M2par"), f"3-D master {n}")
        _same(o.cx(), ks.cx(0), ("sigma", "nu", "poi"), f"collision counts {n}")
        ea, eb = o.ena_master(), ks.ena_master(0)
        assert np.array_equal(ea["nq_ena"], eb["nq_ena"]), n
        assert np.array_equal(ea["esc"], eb["esc"]), n
        _same(o.ena_moments(), ks.ena_moments(0), ("counts", "M0", "M1P", "M1Q", "M1PHI", "M2", "M2P", "M2Q",
                                                   "M2PHI"), f"ENA moments {n}")


@pytest.mark.parametrize("filt", [0, 1])
def test_cx_fp64_bitexact(filt):
    t = _tube(filt=filt)
    ndatfac, nsteps = 2, 12
    print(f"charge exchange fp64: 3-D tube 24 cells, filter {filt}, {nsteps} steps", flush=True)
    o, ks = _pair(t, KAOS_FP64, ndatfac, nsteps)
    for n in range(1, nsteps + 1):
        o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == o.count(), n
        _check_step(o, ks, n, ndatfac)
    assert len(ks.enas(0)["id"]) > 300
    assert ks.cx(0)["poi"].max() > 0
    ks.close()
    o.close()


def test_cx_fp64_two_digit_id_select():
    """40000 ions: ids past 2^14, so the device's id select takes two 14-bit
    digits per cell (one below that); the converted ids are still the first
    nint(poi) of every cell."""
    t = _tube(n_total=40000)
    ndatfac, nsteps = 2, 8
    print(f"charge exchange fp64: 40000 ions, {nsteps} steps", flush=True)
    o, ks = _pair(t, KAOS_FP64, ndatfac, nsteps)
    for n in range(1, nsteps + 1):
        o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == o.count(), n
        _check_step(o, ks, n, ndatfac)
    e = ks.enas(0)["id"]
    assert len(e) > 300 and e.max() >= 2 ** 14
    ks.close()
    o.close()


def test_cx_fp64_enas_escape_with_reorders():
    t = _tube(h=2.0)
    ndatfac, nsteps = 2, 30
    print(f"charge exchange fp64: dt 2 s, {nsteps} steps, reorder every master step", flush=True)
    o, ks = _pair(t, KAOS_FP64, ndatfac, nsteps, reorder=1)
    lb = 0.0
    for n in range(1, nsteps + 1):
        o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == o.count(), n
        _check_step(o, ks, n, ndatfac)
        if n % ndatfac == 0:
            lb += ks.ena_master(0)["esc"][0]
    assert lb > 100
    ks.close()
    o.close()


def test_cx_mixed_close_to_oracle():
    t = _tube()
    ndatfac, nsteps = 2, 12
    print(f"charge exchange mixed: {nsteps} steps", flush=True)
    o, ks = _pair(t, KAOS_MIXED, ndatfac, nsteps)
    for n in range(1, nsteps + 1):
        o.step(n)
        assert ks.step(n).status == 0, n
        if n % ndatfac == 0:
            ea, eb = o.ena_master(), ks.ena_master(0)
            assert np.all(np.abs(ea["nq_ena"] - eb["nq_ena"]) <= 3 + 0.02 * ea["nq_ena"]), n
    na, nb = len(o.enas()["id"]), len(ks.enas(0)["id"])
    assert na > 300 and abs(na - nb) <= 0.02 * na, (na, nb)
    assert abs(o.count() - ks.count(0)) <= 0.02 * o.count()
    ks.close()
    o.close()


def test_cx_refusals(tmp_path):
    t = grids.dipole_tube(nq=24, nvpar=20, nvperp=16, n_total=3000, ratio=1.1)
    src = _tube()
    for k in ("cx", "t_neut", "m_neut", "cx_sigma_scale", "ena_cap", "nvp", "nvq", "nvphi"):
        setattr(t, k, getattr(src, k))
    for k in ("nsneut", "vpGL", "vpGH", "vpGC", "dvp", "vqGL", "vqGH", "vqGC", "dvq", "vphiGL", "vphiGH",
              "vphiGC", "dvphi", "ena_hh", "ena_d33v"):
        t.arrays[k] = src.arrays[k]
    with pytest.raises(KaosError) as e:                     # needs the 3-D grid
        KineticSolver([t], precision=KAOS_FP64)
    assert e.value.code == 1
    t3 = _tube()
    o, ks = _pair(t3, KAOS_FP64, 1, 4)
    o.close()
    for n in range(1, 3):
        assert ks.step(n).status == 0
    with pytest.raises(KaosError) as e:
        ks.save_state(str(tmp_path / "s.bin"))
    assert e.value.code == 4
    ks.close()
