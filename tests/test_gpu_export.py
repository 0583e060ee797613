"""DataExport2Sub's files (kaos_export_master) against the oracle's export
restatement (oracle/export.py), through the C-ABI.

FP64: every file of every exported master step is byte-identical to the one
the oracle writes from its own state (names, sizes, contents), for two tubes
of one context (one southern, one with its own M1Par filter window), with
ndatfac 2.  Mixed: the files hold exactly what the library's own downloads
return.  Errors: export after a non-master step, an unwritable directory,
missing arrays.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import export as oexp
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu


def _tubes():
    a = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=12000, ratio=1.08)
    a.filter_pt_m1par = 5
    a.n_inj_ub = 6
    b = grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=5000, south=True, p=6.0)
    return [a, b]


def _extra(t):
    return np.arccos(t.arrays["cthGC"]), np.linspace(2000.0, 4000.0, t.nq)


def _same_dirs(da, db):
    fa, fb = sorted(os.listdir(da)), sorted(os.listdir(db))
    assert fa == fb, (set(fa) ^ set(fb))
    for n in fa:
        with open(os.path.join(da, n), "rb") as x, open(os.path.join(db, n), "rb") as y:
            if x.read() != y.read():
                raise AssertionError(f"{n} differs")
    return len(fa)


def test_export_fp64_byte_identical_to_oracle(tmp_path):
    tubes = _tubes()
    ndatfac, nsteps, seed = 2, 8, 31
    ions = [grids.initial_ions(t, seed=40 + i) for i, t in enumerate(tubes)]
    caps = [len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 256 for t, (q, _, _) in zip(tubes, ions)]
    orcs = [oracle.Oracle(t, *ions[i], ndatfac=ndatfac, seed=seed, capacity=caps[i], tube_index=i)
            for i, t in enumerate(tubes)]
    ks = KineticSolver(tubes, precision=KAOS_FP64, ndatfac=ndatfac, seed=seed, capacity=caps)
    for i, (q, v, w) in enumerate(ions):
        ks.upload(i, q, v, w)
    print("export fp64: 2 tubes, ndatfac 2, steps", nsteps, flush=True)
    lib_dir, orc_dir = tmp_path / "lib", tmp_path / "orc"
    lib_dir.mkdir()
    orc_dir.mkdir()
    nfiles = 0
    for n in range(1, nsteps + 1):
        for o in orcs:
            o.step(n)
        assert ks.step(n).status == 0
        if n == 1 or n % ndatfac == 0:
            for i, (t, o) in enumerate(zip(tubes, orcs)):
                theta, Ts = _extra(t)
                nn = ks.export_master(i, str(lib_dir), theta, Ts, ns0=1e8, phi=t.phi, species=1, flux_tube=i + 1)
                assert nn == (1 if n == 1 else n // ndatfac + 1)
                assert oexp.export_master(str(orc_dir), t, o.master(), n, ndatfac, theta, Ts, ns0=1e8,
                                          phi=t.phi, s=1, f=i + 1) == nn
            nfiles = _same_dirs(str(lib_dir), str(orc_dir))
    assert nfiles == 5 * (2 * 27 + tubes[0].nq + tubes[1].nq)
    # the filtered M1Par of the first tube used its own window
    m1f = np.fromfile(str(lib_dir / "5_1_1_M1ParFiltAvrgRTfort.bin"))
    assert m1f.any()
    ks.close()


def test_export_3d_grid_fp64_byte_identical_to_oracle(tmp_path):
    # the 3-D branch (ION2VPERPflag = 1): per cell TeNT, N2PerpphRT, d3vCTp in
    # place of NphRT; M1Perp1/M1Perp2 and M2Perp1/M2Perp2 in place of M1Perp, M2Perp
    t = grids.with_vgrid3(grids.dipole_tube(nq=24, nvpar=20, nvperp=16, n_total=8000, ratio=1.1), 9, 7,
                          vperp_max=6.0e3)
    t.arrays["Te"] = np.linspace(8.0e3, 1.2e4, t.nq)          # a per-cell TeNT
    ndatfac, nsteps = 1, 3
    q, v, w = grids.initial_ions(t, seed=12)
    cap = len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 256
    o = oracle.Oracle(t, q, v, w, ndatfac=ndatfac, seed=4, capacity=cap)
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=ndatfac, seed=4, capacity=[cap])
    ks.upload(0, q, v, w)
    rng = np.random.default_rng(3)
    ph = rng.uniform(0.0, 2.0 * np.pi, len(w))
    ks.upload_vperp(0, w * np.cos(ph), w * np.sin(ph))
    o.set_vperp(w * np.cos(ph), w * np.sin(ph))
    print("export 3-D grid fp64: 9x7x20, steps", nsteps, flush=True)
    lib_dir, orc_dir = tmp_path / "lib", tmp_path / "orc"
    lib_dir.mkdir()
    orc_dir.mkdir()
    theta, Ts = _extra(t)
    for n in range(1, nsteps + 1):
        o.step(n)
        assert ks.step(n).status == 0
        nn = ks.export_master(0, str(lib_dir), theta, Ts, ns0=1e8, phi=t.phi)
        assert oexp.export_master(str(orc_dir), t, o.master(), n, ndatfac, theta, Ts, ns0=1e8, phi=t.phi,
                                  master3=o.master3()) == nn
    nfiles = _same_dirs(str(lib_dir), str(orc_dir))
    assert nfiles == (nsteps + 0) * (29 + 3 * t.nq)
    rd = lambda name: np.fromfile(str(lib_dir / f"{nsteps + 1}_1_1_{name}fort.bin"))
    assert rd("5_TeNT").tolist() == [t.arrays["Te"][4]]
    m3 = ks.master3(0)
    got = rd("7_N2PerpphRT").reshape(t.nvpar, t.nv2, t.nv1).T          # (V1, V2, Vpar)
    c = m3["n3_counts"][6]
    assert np.array_equal(got, np.where(c == 0, 0.0, c * (t.nsnormfac / t.arrays["d3xC"][6])))
    assert np.array_equal(rd("M2Perp2phRT"), m3["M2perp2"]) and rd("M1Perp1phRT").any()
    assert not os.path.exists(str(lib_dir / f"{nsteps + 1}_1_1_1_NphRTfort.bin"))
    # the spin-up files of the last master step (SPINUPflag = 1)
    su_lib, su_orc = tmp_path / "su_lib", tmp_path / "su_orc"
    su_lib.mkdir()
    su_orc.mkdir()
    ks.export_spinup(0, str(su_lib), nq_lb=3, nq_ub=t.nq + 2, nsnorm_clb=123.5, nsnorm_cub=0.25,
                     lb_replenish=1, ub_replenish=0, species=2, flux_tube=7)
    oexp.export_spinup(str(su_orc), t, o.master(), o.master3(), 3, t.nq + 2, 123.5, 0.25, 1, 0, s=2, f=7)
    assert _same_dirs(str(su_lib), str(su_orc)) == 11
    ks.close()
    o.close()


def test_export_mixed_holds_the_downloads(tmp_path):
    t = grids.config1(n_total=20000)
    nsteps = 4
    q, v, w = grids.initial_ions(t, seed=2)
    ks = KineticSolver([t], precision=KAOS_MIXED, ndatfac=1, seed=5,
                       capacity=[len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 256])
    ks.upload(0, q, v, w)
    theta, Ts = _extra(t)
    print("export mixed: config 1 tube, steps", nsteps, flush=True)
    for n in range(1, nsteps + 1):
        assert ks.step(n).status == 0
    nn = ks.export_master(0, str(tmp_path) + "/", theta, Ts, ns0=2e8, phi=t.phi)
    assert nn == nsteps + 1
    m = ks.master(0)
    rd = lambda name: np.fromfile(str(tmp_path / f"{nn}_1_1_{name}fort.bin"))
    for name, key in (("M0phRT", "M0"), ("M1PerpphRT", "M1perp"), ("M1ParphRT", "M1par"), ("M2phRT", "M2"),
                      ("M2PerpphRT", "M2perp"), ("M2ParphRT", "M2par"), ("EAPressureRT", "EA"),
                      ("EGmagRT", "EG")):
        assert np.array_equal(rd(name), m[key]), name
    assert rd("TimeT").tolist() == [nsteps * t.h] and rd("ns0GT").tolist() == [2e8]
    for Q in (0, 5, 40):
        got = np.fromfile(str(tmp_path / f"{nn}_1_1_{Q + 1}_NphRTfort.bin")).reshape(t.nvpar, t.nvperp).T
        c = m["nph_counts"][Q]
        want = np.where(c == 0, 0.0, c * (t.nsnormfac / t.arrays["d3xC"][Q]))
        assert np.array_equal(got, want), Q
    ks.close()


def test_export_errors(tmp_path):
    t = grids.dipole_tube(nq=24, nvpar=16, nvperp=16, n_total=2000, ratio=1.1)
    q, v, w = grids.initial_ions(t, seed=1)
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=3, seed=1, capacity=[len(q) + 512])
    ks.upload(0, q, v, w)
    theta, Ts = _extra(t)
    with pytest.raises(KaosError):                     # nothing stepped yet
        ks.export_master(0, str(tmp_path), theta, Ts, ns0=1e8, phi=t.phi)
    assert ks.step(1).status == 0
    assert ks.export_master(0, str(tmp_path), theta, Ts, ns0=1e8, phi=t.phi) == 1
    assert ks.step(2).status == 0                      # not a master step (ndatfac 3)
    with pytest.raises(KaosError):
        ks.export_master(0, str(tmp_path), theta, Ts, ns0=1e8, phi=t.phi)
    assert ks.step(3).status == 0
    with pytest.raises(KaosError) as e:
        ks.export_master(0, str(tmp_path / "no" / "such" / "dir"), theta, Ts, ns0=1e8, phi=t.phi)
    assert e.value.code == 6
    with pytest.raises(ValueError):
        ks.export_master(0, str(tmp_path), theta[:-1], Ts, ns0=1e8, phi=t.phi)
    assert ks.export_master(0, str(tmp_path), theta, Ts, ns0=1e8, phi=t.phi) == 2
    with pytest.raises(KaosError) as e:                 # spin-up files need the 3-D moments
        ks.export_spinup(0, str(tmp_path))
    assert e.value.code == 4
    ks.close()
