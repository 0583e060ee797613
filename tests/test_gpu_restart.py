"""Restart from spin-up files (SPINUPflag = 0, _1_SimParameterization.f90:801-946)
through the C-ABI, FP64 bit-exact against the oracle.

A 3-D grid run on the device writes its spin-up files (kaos_export_spinup)
after its last master step.  kaos_read_spinup reads them back; the restart
then takes, as the reference's restart does:
* the per-cell macro-ion counts from DensityOutputRT (kaos_cell_counts, the
  nsnormC0 of _0_1_1_ConfigGridGeneratorB.f90:75-86);
* the per-cell temperatures of the initial Maxwellians from
  TemperatureOutputRT (kaos_seed_particles, _6_VelocityDistribution.f90:70-73);
* the LB injection count nint(nsnormCLBGT) (one rank, BC:747-750), where
  nsnormCLBGT = nsnormC0(NqLB + 1) is the first computational cell's count
  (_0_1_1_ConfigGridGeneratorB.f90:189; nsnormCLBInput fills only the ghost
  cell nsnormC0(NqLB), :76-78);
* the ambipolar field EAmagOutputRT (kaos_set_field), at step 1 only or at
  every master step (_8_7_AmbipolarEfield.f90:1375-1407).
The oracle restarts from the same arrays (orc_seed_ions, set_field_input);
every per-ion value and every master-step output is bit-exact for 6 steps.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from kaos_amd import grids, _abi
from kaos_amd.solver import (KineticSolver, KaosError, read_spinup, KAOS_FP64, KAOS_FIELD_INPUT_FIRST,
                             KAOS_FIELD_INPUT_FIXED, KAOS_FIELD_COMPUTED)

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
MASTER_KEYS = ("nq_counts", "NqRT", "M0", "M1par", "M2", "M2par", "EA", "EG", "esc")


def _tube():
    t = grids.with_vgrid3(grids.dipole_tube(nq=24, nvpar=20, nvperp=16, n_total=12000, ratio=1.1), 10, 8,
                          vperp_max=6.0e3)
    return t


def _spinup_dir(tmp_path, t):
    """A 3-D FP64 run of 4 steps on the device and its spin-up files."""
    q, v, w = grids.initial_ions(t, seed=31)
    cap = len(q) + 8 * (t.n_inj_lb + t.n_inj_ub) + 256
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=1, seed=8, capacity=[cap])
    ks.upload(0, q, v, w)
    for n in range(1, 5):
        assert ks.step(n).status == 0
    d = tmp_path / "spinup"
    d.mkdir()
    ks.export_spinup(0, str(d), nq_lb=1, nq_ub=t.nq, nsnorm_clb=37.4, nsnorm_cub=0.0)
    ks.close()
    return d


@pytest.mark.parametrize("mode", [KAOS_FIELD_INPUT_FIRST, KAOS_FIELD_INPUT_FIXED])
def test_restart_from_spinup_files_bitexact(tmp_path, mode):
    t = _tube()
    print(f"restart: 3-D tube 24 cells, field mode {mode}", flush=True)
    sp = read_spinup(_spinup_dir(tmp_path, t))
    # the upper cells of this exponential profile hold no ions: density and temperature 0 there
    pop = sp["density"] > 0
    assert sp["nq"] == t.nq and pop.sum() >= t.nq // 2 and np.all(sp["temperature"][pop] > 0)
    assert np.all(sp["density"][~pop] == 0) and sp["nsnorm_clb"] == 37.4
    counts = np.zeros(t.nq, np.int64)
    d3x = np.ascontiguousarray(t.arrays["d3xC"])
    assert _abi.load().kaos_cell_counts(t.nq, _abi.ptr(sp["density"]), _abi.ptr(d3x), t.nsnormfac,
                                        _abi.ptr(counts, C.c_int64)) == 0
    assert counts.sum() > 1000
    t.n_inj_lb = int(counts[0])                              # nsnormCLBGT: the first cell's count, one rank
    t.Ts_lb = float(sp["temperature"][0])
    temp = sp["temperature"]
    nsteps, seed = 6, 4242
    cap = int(counts.sum()) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 256
    s = oracle.seed_ions(t, counts, temp, seed)
    o = oracle.Oracle(t, s["q0"], s["vpar"], s["vperp"], ndatfac=1, seed=seed, capacity=cap)
    o.set_vperp(s["v1"], s["v2"])
    o.set_field_input(sp["ea_mag"], mode)
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=1, seed=seed, capacity=[cap])
    if mode == KAOS_FIELD_INPUT_FIRST:          # set before the device state exists ...
        ks.set_field(0, sp["ea_mag"], mode)
    ks.seed(0, counts, temp, seed)
    if mode == KAOS_FIELD_INPUT_FIXED:          # ... and after it
        ks.set_field(0, sp["ea_mag"], mode)
    for n in range(1, nsteps + 1):
        o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == o.count(), n
        a, b = o.particles(), ks.particles(0)
        for k in PART_KEYS:
            assert np.array_equal(a[k], b[k]), (n, k)
        ma, mb = o.master(), ks.master(0)
        for k in MASTER_KEYS:
            assert np.array_equal(ma[k], mb[k]), (n, k)
        m3a, m3b = o.master3(), ks.master3(0)
        assert np.array_equal(m3a["n3_counts"], m3b["n3_counts"]), n
        if n == 1 or mode == KAOS_FIELD_INPUT_FIXED:
            assert np.array_equal(mb["EA"], sp["ea_mag"]), n
        else:
            assert not np.array_equal(mb["EA"], sp["ea_mag"]), n
    ks.close()
    o.close()


def test_set_field_between_steps_and_errors():
    t = grids.config1(n_total=5000)
    q, v, w = grids.initial_ions(t, seed=2)
    cap = len(q) + 12 * (t.n_inj_lb + t.n_inj_ub) + 256
    o = oracle.Oracle(t, q, v, w, ndatfac=2, seed=6, capacity=cap)
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=2, seed=6, capacity=[cap])
    ks.upload(0, q, v, w)
    ea = np.linspace(1e-7, -1e-7, t.nq)
    print("set_field between steps: config 1 tube, 5000 ions", flush=True)
    for n in range(1, 9):
        if n == 3:          # a fixed field from master step 4 on, set after graphs were built
            ks.set_field(0, ea, KAOS_FIELD_INPUT_FIXED)
            o.set_field_input(ea, 2)
        if n == 7:          # back to the computed field
            ks.set_field(0, None, KAOS_FIELD_COMPUTED)
            o.set_field_input(None, 0)
        o.step(n)
        assert ks.step(n).status == 0
        a, b = o.particles(), ks.particles(0)
        for k in PART_KEYS:
            assert np.array_equal(a[k], b[k]), (n, k)
        if n % 2 == 0:
            assert np.array_equal(o.master()["EA"], ks.master(0)["EA"]), n
    assert np.array_equal(ks.master(0)["EA"], o.master()["EA"])
    for bad in ((ea, 3), (None, KAOS_FIELD_INPUT_FIXED), (np.full(t.nq, np.nan), KAOS_FIELD_INPUT_FIRST)):
        with pytest.raises(KaosError) as e:
            ks.set_field(0, *bad)
        assert e.value.code == 1
    with pytest.raises(KaosError):
        ks.set_field(5, ea, KAOS_FIELD_INPUT_FIRST)
    ks.close()
    o.close()
