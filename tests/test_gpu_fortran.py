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
