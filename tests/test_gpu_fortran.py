"""The Fortran driver (kaos_amd/fortran/kaos_fdriver, iso_c_binding interface
kaos_hip_iface.f90) runs the hot path through the C-ABI and reproduces the
oracle bit for bit in the FP64 form: counts, moments, field and escapes of
every master step."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd import fortran as kf

pytestmark = pytest.mark.gpu

MOMENTS = ("M0", "M1perp", "M1par", "M2", "M2perp", "M2par")


@pytest.mark.parametrize("ndatfac", [1, 3])
def test_fortran_driver_matches_oracle(tmp_path, ndatfac):
    if not os.path.exists(kf.DRIVER):
        pytest.fail("kaos_fdriver is not built (kaos_amd.build.build_fortran)")
    tube = grids.config1(n_total=20000)
    q0, v0, w0 = grids.initial_ions(tube, seed=4)
    nsteps, seed = 10, 21
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    kf.write_input(fin, tube, q0, v0, w0, nsteps, ndatfac=ndatfac, seed=seed)
    print(f"[fortran] running kaos_fdriver ndatfac={ndatfac}", flush=True)
    r = subprocess.run([kf.DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    recs = kf.read_output(fout, tube.nq)
    orc = oracle.Oracle(tube, q0, v0, w0, ndatfac=ndatfac, seed=seed,
                        capacity=len(q0) + (nsteps + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 1024)
    masters = [n for n in range(1, nsteps + 1) if n == 1 or n % ndatfac == 0]
    assert [x["n"] for x in recs] == masters
    it = iter(recs)
    for n in range(1, nsteps + 1):
        orc.step(n)
        if n not in masters:
            continue
        rec = next(it)
        m = orc.master()
        assert rec["n_alive"] == orc.count()
        assert np.array_equal(rec["nq_counts"], m["nq_counts"]), n
        for i, k in enumerate(MOMENTS):
            assert np.array_equal(rec["M"][i], m[k]), (n, k)
        assert np.array_equal(rec["EA"], m["EA"]), n
        assert np.array_equal(rec["esc"], m["esc"]), n
    orc.close()


def test_fortran_driver_seed_and_export(tmp_path):
    """The Fortran driver seeds the ions on the device (kaos_cell_counts +
    kaos_seed_particles) and exports every master step (kaos_export_master):
    the files equal, byte for byte, the oracle's export of the oracle started
    from its own restatement of the seeding."""
    from oracle import export as oexp
    if not os.path.exists(kf.DRIVER):
        pytest.fail("kaos_fdriver is not built (kaos_amd.build.build_fortran)")
    tube = grids.dipole_tube(nq=48, nvpar=32, nvperp=32, n_total=15000, ratio=1.06)
    nsteps, ndatfac, seed, iseed = 6, 2, 9, 777
    lib = oracle.load()
    density = 1e10 * np.exp(-tube.s_centre / 4.0e5)
    counts = np.array([lib.orc_cell_count(a, b, tube.nsnormfac) for a, b in zip(density, tube.arrays["d3xC"])])
    ex = dict(theta=np.arccos(tube.arrays["cthGC"]), Ts=np.linspace(2500.0, 5000.0, tube.nq), density=density,
              phi=tube.phi, ns0=1e8)
    cap = int(counts.sum()) + (nsteps + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 512
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    fdir, odir = tmp_path / "fexp", tmp_path / "oexp"
    fdir.mkdir()
    odir.mkdir()
    kf.write_input(fin, tube, np.zeros(0), np.zeros(0), np.zeros(0), nsteps, ndatfac=ndatfac, seed=seed,
                   capacity=cap, export=ex)
    print(f"[fortran] seeded run with export, ions {int(counts.sum())}", flush=True)
    r = subprocess.run([kf.DRIVER, fin, fout, str(fdir), str(iseed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s = oracle.seed_ions(tube, counts, ex["Ts"], iseed)
    orc = oracle.Oracle(tube, s["q0"], s["vpar"], s["vperp"], ndatfac=ndatfac, seed=seed, capacity=cap)
    orc.set_vperp(s["v1"], s["v2"])
    for n in range(1, nsteps + 1):
        orc.step(n)
        if n == 1 or n % ndatfac == 0:
            oexp.export_master(str(odir), tube, orc.master(), n, ndatfac, ex["theta"], ex["Ts"], ns0=ex["ns0"],
                               phi=ex["phi"])
    orc.close()
    names = sorted(os.listdir(fdir))
    assert names == sorted(os.listdir(odir)) and len(names) == 4 * (27 + tube.nq)
    for nm in names:
        assert (fdir / nm).read_bytes() == (odir / nm).read_bytes(), nm


def test_fortran_driver_3d_grid_seed_and_export(tmp_path):
    """The Fortran driver on the 3-D velocity grid (GRID3D: KAOS_FLAG_VPERP3D
    and the kaos_grid 3-D fields through iso_c_binding), ions seeded on the
    device: the eight 3-D moments of every master step (kaos_download_master3d)
    equal the oracle's bit for bit, and the export directory (the 3-D
    branch's files and the spin-up files) equals the oracle's byte for byte."""
    from oracle import export as oexp
    if not os.path.exists(kf.DRIVER):
        pytest.fail("kaos_fdriver is not built (kaos_amd.build.build_fortran)")
    tube = grids.with_vgrid3(grids.dipole_tube(nq=32, nvpar=24, nvperp=16, n_total=8000, ratio=1.08), 12, 10,
                             vperp_max=6.0e3)
    nsteps, ndatfac, seed, iseed = 6, 2, 4, 555
    lib = oracle.load()
    density = 1e10 * np.exp(-tube.s_centre / 4.0e5)
    counts = np.array([lib.orc_cell_count(a, b, tube.nsnormfac) for a, b in zip(density, tube.arrays["d3xC"])])
    ex = dict(theta=np.arccos(tube.arrays["cthGC"]), Ts=np.linspace(2500.0, 5000.0, tube.nq), density=density,
              phi=tube.phi, ns0=1e8)
    cap = int(counts.sum()) + (nsteps + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 512
    fin, fout, fg3 = str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(tmp_path / "grid3.bin")
    fdir, odir = tmp_path / "fexp", tmp_path / "oexp"
    fdir.mkdir()
    odir.mkdir()
    kf.write_input(fin, tube, np.zeros(0), np.zeros(0), np.zeros(0), nsteps, ndatfac=ndatfac, seed=seed,
                   capacity=cap, export=ex)
    kf.write_grid3(fg3, tube)
    print(f"[fortran] 3-D grid seeded run with export, ions {int(counts.sum())}", flush=True)
    r = subprocess.run([kf.DRIVER, fin, fout, str(fdir), str(iseed), fg3], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    recs = kf.read_output(fout, tube.nq, v3d=True)
    s = oracle.seed_ions(tube, counts, ex["Ts"], iseed)
    orc = oracle.Oracle(tube, s["q0"], s["vpar"], s["vperp"], ndatfac=ndatfac, seed=seed, capacity=cap)
    orc.set_vperp(s["v1"], s["v2"])
    k = 0
    for n in range(1, nsteps + 1):
        orc.step(n)
        if n == 1 or n % ndatfac == 0:
            m, m3 = orc.master(), orc.master3()
            assert recs[k]["n"] == n
            for i, name in enumerate(oracle.MOMENTS3):
                assert np.array_equal(recs[k]["M8"][i], m3[name]), (n, name)
            assert np.array_equal(recs[k]["nq_counts"], m["nq_counts"]), n
            oexp.export_master(str(odir), tube, m, n, ndatfac, ex["theta"], ex["Ts"], ns0=ex["ns0"],
                               phi=ex["phi"], master3=m3)
            oexp.export_spinup(str(odir), tube, m, m3, 1, tube.nq, float(tube.n_inj_lb), float(tube.n_inj_ub),
                               0, 0)
            k += 1
    orc.close()
    assert k == len(recs) == 4
    names = sorted(os.listdir(fdir))
    assert names == sorted(os.listdir(odir)) and len(names) == 4 * (29 + 3 * tube.nq) + 11
    for nm in names:
        assert (fdir / nm).read_bytes() == (odir / nm).read_bytes(), nm
