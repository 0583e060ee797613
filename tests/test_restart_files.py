"""Restart input (SPINUPflag = 0): the spin-up files read back through the
C-ABI (kaos_read_spinup; host file reads, no device) and the oracle's input
field, on the CPU.

* Round trip: files written by the oracle's restatement of DataExport2Sub's
  spin-up block (oracle/export.py, whose byte layout is pinned against flang's
  runtime in test_oracle_export.py) read back by kaos_read_spinup equal the
  arrays and scalars written, for two (species, flux tube) pairs in one
  directory and with the reference's ghost-cell bounds (NqLB = 3).
* Refusals: a missing file, a short or long array file, NqUB < NqLB, and more
  cells than the caller's arrays hold.
* The oracle's input field (_8_7_AmbipolarEfield.f90:1375-1407): in the fixed
  mode EAmagRT equals the input at step 1 and at every master step; in the
  first-step mode it equals the input at step 1 only and is afterwards the
  stencil field (orc_ambipolar) of the previous master step's density; the
  ions' step-1 AEAmag come from the input (doubling the input doubles every
  one of them), not from the computed field.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from oracle import export as oexp
from kaos_amd import grids
from kaos_amd.solver import read_spinup, KaosError


def _write(directory, nq, rng, s, f, nq_lb=3):
    m3 = {k: rng.uniform(1.0, 2.0, nq) * sc for k, sc in
          (("M0", 1e9), ("M2par", 1e-20), ("M2perp1", 2e-20), ("M2perp2", 3e-20))}
    master = {"EA": rng.normal(0.0, 1e-6, nq)}
    t = type("T", (), {"nq": nq})()
    T = oexp.export_spinup(str(directory), t, master, m3, nq_lb, nq_lb + nq - 1, 4321.25, 0.5, 1, 0,
                           s=s, f=f)
    return m3, master, T


def test_read_spinup_round_trip(tmp_path):
    rng = np.random.default_rng(11)
    m3a, ma, Ta = _write(tmp_path, 37, rng, 1, 1)
    m3b, mb, Tb = _write(tmp_path, 64, rng, 2, 5)
    a = read_spinup(tmp_path, 1, 1)
    b = read_spinup(tmp_path, 2, 5)
    for d, m3, m, T, nq in ((a, m3a, ma, Ta, 37), (b, m3b, mb, Tb, 64)):
        assert (d["nq"], d["nq_lb"], d["nq_ub"]) == (nq, 3, nq + 2)
        assert (d["nsnorm_clb"], d["nsnorm_cub"], d["lb_replenish"], d["ub_replenish"]) == (4321.25, 0.5, 1, 0)
        assert np.array_equal(d["density"], m3["M0"])
        assert np.array_equal(d["temperature"], T)
        assert np.array_equal(d["ea_inertial"], np.zeros(nq))
        assert np.array_equal(d["ea_pressure"], m["EA"])
        assert np.array_equal(d["ea_mag"], m["EA"])
    # the temperature written is the reference's (2/3)(M2par + M2perp1 + M2perp2)/kB, left to right
    c = 2.0 / 3.0
    want = ((c * m3a["M2par"]) / 1.38e-23 + (c * m3a["M2perp1"]) / 1.38e-23) + (c * m3a["M2perp2"]) / 1.38e-23
    assert np.array_equal(a["temperature"], want)


def test_read_spinup_refusals(tmp_path):
    rng = np.random.default_rng(12)
    with pytest.raises(KaosError) as e:                     # nothing there
        read_spinup(tmp_path, 1, 1)
    assert e.value.code == 6 and "NqLBT" in str(e.value)
    _write(tmp_path, 20, rng, 1, 1)
    with pytest.raises(KaosError) as e:                     # more cells than the arrays hold
        read_spinup(tmp_path, 1, 1, nq_cap=19)
    assert e.value.code == 1
    assert read_spinup(tmp_path, 1, 1, nq_cap=20)["nq"] == 20
    p = os.path.join(tmp_path, "1_1_TemperatureOutputRTfort.bin")
    full = open(p, "rb").read()
    for bad in (full[:-8], full + b"\0"):                   # short, then one byte long
        open(p, "wb").write(bad)
        with pytest.raises(KaosError) as e:
            read_spinup(tmp_path, 1, 1)
        assert e.value.code == 6 and "TemperatureOutputRT" in str(e.value)
    open(p, "wb").write(full)
    np.array([30], "<i8").tofile(os.path.join(tmp_path, "1_1_NqLBTfort.bin"))   # NqUB < NqLB
    with pytest.raises(KaosError) as e:
        read_spinup(tmp_path, 1, 1)
    assert e.value.code == 6


def _orc(t, ndatfac, seed=3):
    q, v, w = grids.initial_ions(t, seed=21)
    return oracle.Oracle(t, q, v, w, ndatfac=ndatfac, seed=seed, capacity=len(q) + 40 * (t.n_inj_lb + 8) + 64)


def test_oracle_field_input_modes():
    t = grids.config1(n_total=6000)
    ea_in = np.linspace(-2e-7, 3e-7, t.nq)
    ndatfac = 2
    fixed, first, plain, twice = (_orc(t, ndatfac) for _ in range(4))
    fixed.set_field_input(ea_in, 2)
    first.set_field_input(ea_in, 1)
    twice.set_field_input(2.0 * ea_in, 1)
    for o in (fixed, first, plain, twice):
        o.step(1)
    assert np.array_equal(fixed.master()["EA"], ea_in) and np.array_equal(first.master()["EA"], ea_in)
    assert not np.array_equal(plain.master()["EA"], ea_in)
    # the ions see the input, not the computed field: with the input doubled
    # every AEAmag of step 1 doubles (interpolation is linear in the field)
    a1, a2 = first.particles()["aea"], twice.particles()["aea"]
    assert np.all(a1 != 0.0) and np.allclose(a2, 2.0 * a1, rtol=1e-15, atol=0.0)
    cfg = oracle.cfg_of(t, ndatfac, 3)
    lib = oracle.load()
    prev_m0 = first.master()["M0"].copy()
    for n in range(2, 9):
        for o in (fixed, first, plain, twice):
            o.step(n)
        if n % ndatfac == 0:
            assert np.array_equal(fixed.master()["EA"], ea_in), n
            # first-step mode: the stencil field of the previous master step's density
            want, m0f = np.zeros(t.nq), np.zeros(t.nq)
            lib.orc_ambipolar(C.byref(cfg), oracle.ptr(prev_m0), oracle.ptr(m0f), oracle.ptr(want))
            got = first.master()
            assert np.array_equal(got["EA"], want), n
            prev_m0 = got["M0"].copy()
    for o in (fixed, first, plain, twice):
        o.close()
