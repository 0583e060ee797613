"""Pins of the oracle's export restatement (oracle/export.py) of
DataExport2Sub (_8_0_2_DataExport2.f90:25-704), CPU only:
- names and byte layout against what the Fortran runtime itself writes for
  the reference's statements (oracle/export_probe.f90, compiled with flang);
- the files read back the way the reference's reader does
  (KAOS_Data_Plots.py:60-115: str(nn) + '_1_1_' + name, numpy.fromfile);
- M1ParFiltAvrgRT = the moving-average filter of the previous master step's
  M1Par with its own window (_8_7_AmbipolarEfield.f90:64-80).
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from oracle import export as oexp
from oracle import ptr
from kaos_amd import grids

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("flang not present")
    d = tmp_path_factory.mktemp("probe")
    exe = str(d / "probe")
    r = subprocess.run([FLANG, os.path.join(ROOT, "oracle", "export_probe.f90"), "-o", exe],
                       capture_output=True, text=True, cwd=str(d))
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("nn,s,f,Q", [(1, 1, 1, 1), (123, 1, 2, 45), (99999, 12, 345, 7), (10, 3, 1, 256)])
def test_names_and_stream_bytes_match_the_fortran_runtime(probe, tmp_path, nn, s, f, Q):
    r = subprocess.run([probe, str(nn), str(s), str(f), str(Q)], capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    names = r.stdout.split()
    assert names == [oexp.file_name(nn, s, f, "TimeT"), oexp.file_name(nn, s, f, "M1Perp2FiltAvrgRT"),
                     oexp.file_name(nn, s, f, "NphRT", Q), oexp.spinup_name(s, f, "TemperatureOutputRT")]
    # stream writes: no record markers, Fortran (column-major) order, int64
    got = {n: (tmp_path / n).read_bytes()
           for n in ("probe_scalar.bin", "probe_int.bin", "probe_array.bin", "probe_array3.bin")}
    mine = tmp_path / "mine"
    mine.mkdir()
    oexp._put(str(mine / "x"), [0.1 * nn])
    oexp._put(str(mine / "k"), [7 * nn], "<i8")
    a = np.array([[10 * i + j + 0.25 for j in range(1, 5)] for i in range(1, 4)])   # a(i, j)
    oexp._put(str(mine / "a"), np.asfortranarray(a).ravel(order="F"))
    assert (mine / "x").read_bytes() == got["probe_scalar.bin"]
    assert (mine / "k").read_bytes() == got["probe_int.bin"]
    assert (mine / "a").read_bytes() == got["probe_array.bin"]
    # 3-D: b(i, j, l) = 100 i + 10 j + l + 1/2 written whole is the layout of
    # N2PerpphRTp(:, :, :, nn) and d3vCTp; n2perp_rt of counts [Q][V1][V2][Vpar]
    # = 100 i + 10 j + l (weight 1) must give the same sequence less 1/2
    b3 = np.frombuffer(got["probe_array3.bin"], "<f8")
    cnt = np.array([[[100 * i + 10 * j + l for l in range(1, 5)] for j in range(1, 4)] for i in range(1, 3)])
    assert np.array_equal(oexp.n2perp_rt(cnt[None], 1.0, np.ones(1))[0], b3 - 0.5)


def _run(tube, nsteps, ndatfac=1, seed=3):
    q0, v0, w0 = grids.initial_ions(tube, seed=8)
    o = oracle.Oracle(tube, q0, v0, w0, ndatfac=ndatfac, seed=seed,
                      capacity=len(q0) + (nsteps + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 64)
    return o


def test_export_reads_back_with_the_reference_reader(tmp_path):
    t = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=6000, ratio=1.08)
    o = _run(t, 4)
    theta = np.arccos(t.arrays["cthGC"])
    Ts = np.full(t.nq, t.Ts_lb)
    for n in range(1, 5):
        o.step(n)
    m = o.master()
    nn = oexp.export_master(str(tmp_path), t, m, 4, 1, theta, Ts, ns0=1e8, phi=t.phi)
    assert nn == 5
    datadir = str(tmp_path) + "/"
    rd = lambda name: np.fromfile(datadir + str(nn) + "_" + str(1) + "_" + str(1) + "_" + name + "fort.bin")
    assert rd("TimeT").tolist() == [4 * t.h]
    assert np.fromfile(datadir + "5_1_1_ndatfacGTfort.bin", dtype=np.int64).tolist() == [1]
    for name, want in (("M0phRT", m["M0"]), ("M1PerpphRT", m["M1perp"]), ("M1ParphRT", m["M1par"]),
                       ("M2phRT", m["M2"]), ("M2PerpphRT", m["M2perp"]), ("M2ParphRT", m["M2par"]),
                       ("EAmagRT", m["EA"]), ("EGmagRT", m["EG"]), ("M0FiltAvrgRT", m["M0F"]),
                       ("rGCGT", t.arrays["rGC"]), ("pGCGT", np.full(t.nq, t.p)), ("thetaGCGT", theta)):
        assert np.array_equal(rd(name), want), name
    for name in ("EPmagRT", "PhiParRT", "EAInertialRT", "M2ParFiltAvrgRT"):
        assert not rd(name).any(), name
    # the distribution: counts x nsnormfac / d3x, (Vperp, Vpar) in Fortran order
    for Q in (0, 3, 17):
        v = np.fromfile(datadir + f"{nn}_1_1_{Q + 1}_NphRTfort.bin").reshape(t.nvpar, t.nvperp).T
        c = m["nph_counts"][Q]
        assert np.array_equal(v == 0, c == 0)
        assert np.allclose(v[c > 0], c[c > 0] * t.nsnormfac / t.arrays["d3xC"][Q], rtol=1e-15, atol=0)
    assert len(os.listdir(tmp_path)) == 27 + t.nq
    o.close()


@pytest.mark.parametrize("w1", [0, 5])
def test_m1par_filter_uses_the_previous_master_step(w1):
    t = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=6000, ratio=1.08)
    t.filter_pt_m1par = w1
    o = _run(t, 6, ndatfac=2)
    lib = oracle.load()
    o.step(1)
    prev = o.master()
    assert not prev["M1F"].any()                 # no master step before the first
    for n in range(2, 7):
        o.step(n)
        if n % 2:
            continue
        m = o.master()
        want = np.zeros(t.nq)
        lib.orc_filter(ptr(np.ascontiguousarray(prev["M1par"])), ptr(want), t.nq, w1 or t.filter_pt)
        assert np.array_equal(m["M1F"], want), n
        other = np.zeros(t.nq)
        lib.orc_filter(ptr(np.ascontiguousarray(prev["M1par"])), ptr(other), t.nq, 7)
        assert not np.array_equal(m["M1F"], other)
        prev = m
    o.close()


def test_spinup_temperature_of_a_maxwellian(tmp_path):
    # TemperatureOutputRT = (2/3)(M2Par + M2Perp1 + M2Perp2)/kB of ions seeded
    # from a Maxwellian at temperature T0 is T0 up to the binning of the
    # velocity grid and sampling (a dropped term gives 2/3 T0, a wrong factor
    # 3/2 or 1/3 misses by 50% or more); DensityOutputRT is M0
    t = grids.with_vgrid3(grids.dipole_tube(nq=16, nvpar=40, nvperp=20, n_total=2000, ratio=1.1), 40, 40,
                          vperp_max=1.2e4)
    T0 = 5000.0
    counts = np.full(t.nq, 4000, np.int64)
    ions = oracle.seed_ions(t, counts, np.full(t.nq, T0), seed=7)
    o = oracle.Oracle(t, ions["q0"], ions["vpar"], ions["vperp"], capacity=len(ions["q0"]) + 100)
    o.set_vperp(ions["v1"], ions["v2"])
    o.step(1)
    m, m3 = o.master(), o.master3()
    T = oexp.export_spinup(str(tmp_path), t, m, m3, 1, t.nq, 3.0, 4.0, 0, 1)
    assert np.all(np.abs(T / T0 - 1.0) < 0.05), T / T0
    rd = lambda name, dt="<f8": np.fromfile(str(tmp_path / oexp.spinup_name(1, 1, name)), dtype=dt)
    assert np.array_equal(rd("TemperatureOutputRT"), T) and np.array_equal(rd("DensityOutputRT"), m3["M0"])
    assert rd("NqUBT", "<i8").tolist() == [t.nq] and rd("UBREPLENISHflagTOutput", "<i8").tolist() == [1]
    assert len(os.listdir(tmp_path)) == 11
    o.close()
