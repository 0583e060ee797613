"""Grid set-up on the device (kaos_generate_grid) through the C-ABI.

Bit-exact against the oracle's set-up (oracle.generate_grid, the same
arithmetic on the host; pinned against the reference's own set-up in
test_oracle_gridgen.py) for: the reference harness's tube (3-D grid,
symmetric V||), a southern tube on the 2-D grid with the one-sided V|| grid,
and a 256-cell tube with charge exchange (neutral counts, the ENA grid and its
metric per config cell and bin).  Then a tube built from the generated grids
(grids.tube_from_setup) with its ions seeded on the device from the
generated counts runs 6 FP64 steps bit-exact against the oracle.
"""
import math

import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64, generate_grid

pytestmark = pytest.mark.gpu

RE, MO = 6.38e6, 16.0 * 1.66054e-27
HARNESS = dict(nq_ub=64, qGA=0.7790, qGB=0.0556, L=4.0, phi=math.pi / 2.0, ms=MO, ti=3482.6, te=1e4, ns0=1e8,
               nsnormfac=3.85e17, vperp3d=1, symvpar=1, cx=0, nvpar_gp=14, nvperp_gp=14, vpar_sigma_fac=7.4,
               vperp_sigma_fac=7.4)
CASES = {
    "harness": HARNESS,
    "south_2d_onesided": dict(HARNESS, qGA=-0.7790, qGB=-0.0556, vperp3d=0, symvpar=0, nvpar_gp=40, nvperp_gp=30),
    "cx_256": dict(HARNESS, nq_ub=256, L=6.0, qGA=0.5, qGB=0.01, cx=1, t_neut=848.0, zns0_neut=RE + 719e3,
                   ns0_neut=1.423e11, nvpar_gp=20, nvperp_gp=12),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_generated_grid_equals_the_oracle(case):
    p = CASES[case]
    print(f"grid set-up on the device: {case}", flush=True)
    a, b = generate_grid(p), oracle.generate_grid(p)
    for k in ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Ts", "Te", "density", "counts"):
        assert np.array_equal(a[k], b[k]), (case, k)
    if p["cx"]:
        assert np.array_equal(a["ns_neut"], b["ns_neut"])
        assert np.array_equal(a["ena_hh"], b["ena_hh"]) and np.array_equal(a["ena_d33v"], b["ena_d33v"])
    for ax, n in b["ncell"].items():
        if n:
            for x, y in zip(a[ax], b[ax]):
                assert np.array_equal(x, y), (case, ax)
    k3 = "d3v3" if p["vperp3d"] else "d3v"
    assert np.array_equal(a[k3], b[k3])
    assert a["nsnorm_clb"] == b["counts"][0] and a["nsnorm_cub"] == b["counts"][-1]


def test_run_on_generated_grids_bitexact():
    p = HARNESS
    g = generate_grid(p)
    t = grids.tube_from_setup(g, p, h=0.05, qs=1.602e-19, ranks=1)
    print(f"6 fp64 steps on the generated grid: {int(t.cell_counts.sum())} seeded ions", flush=True)
    temp = np.asarray(g["Ts"])
    s = oracle.seed_ions(t, t.cell_counts, temp, 77)
    nsteps, ndatfac = 6, 3
    cap = int(t.cell_counts.sum()) + (nsteps + 2) * t.n_inj_lb + 256
    o = oracle.Oracle(t, s["q0"], s["vpar"], s["vperp"], ndatfac=ndatfac, seed=3, capacity=cap)
    o.set_vperp(s["v1"], s["v2"])
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=ndatfac, seed=3, capacity=[cap])
    ks.seed(0, t.cell_counts, temp, 77)
    for n in range(1, nsteps + 1):
        o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == o.count(), n
        pa, pb = o.particles(), ks.particles(0)
        for k in ("x", "y", "z", "Vx", "Vy", "Vz", "qind", "vparind", "vperpind"):
            assert np.array_equal(pa[k], pb[k]), (n, k)
        if n == 1 or n % ndatfac == 0:
            ma, mb = o.master3(), ks.master3(0)
            for k in ("n3_counts", "M0", "M1par", "M2"):
                assert np.array_equal(ma[k], mb[k]), (n, k)
    ks.close()
    o.close()
