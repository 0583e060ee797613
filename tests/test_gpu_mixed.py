"""Mixed-precision form of the HIP path against the FP64 oracle, through the C-ABI.

The mixed form keeps binary32 stage arithmetic and 32-bit state words, so it
is not bit-exact; the bar is closeness to the oracle on the same inputs:
  * the same ions alive after every step (removal and injection semantics);
  * integer counts equal but for a few ions whose binary32 trajectory crossed
    a cell edge on the other side of the binary64 one;
  * positions within 1e-9 and velocities within 1e-6 of the largest value
    (the loosest are ions off the Vpar grid, whose V|| is stored as a
    binary32 offset from the grid edge);
  * moments and the ambipolar field within the relative L2 errors below
    (per-step, over the config cells).
The headline size (1e8 ions) is measured by bench.py, which reports the
moment errors of the mixed form against the FP64 form there.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_MIXED

pytestmark = pytest.mark.gpu

MOMENTS = ("M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA")


def _l2rel(a, b):
    n = np.linalg.norm(a)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a - b)


def _run(tube, nsteps, ndatfac=1, uniforms=False, ions_seed=3, seed=11, max_dcount=2, tol_l2=1e-4,
         reorder_every=0):
    q0, v0, w0 = grids.initial_ions(tube, seed=ions_seed)
    cap = len(q0) + 20000
    orc = oracle.Oracle(tube, q0, v0, w0, ndatfac=ndatfac, seed=seed, capacity=cap)
    ks = KineticSolver([tube], precision=KAOS_MIXED, ndatfac=ndatfac, seed=seed, capacity=[cap],
                       reorder_every=reorder_every)
    ks.upload(0, q0, v0, w0)
    rng = np.random.default_rng(99)
    ninj = (tube.n_inj_lb if tube.lbcond else 0) + (tube.n_inj_ub if tube.ubcond else 0)
    worst = {k: 0.0 for k in MOMENTS}
    for n in range(1, nsteps + 1):
        master = n >= 2 and n % ndatfac == 0
        u = 1.0 - rng.random(6 * ninj) if (uniforms and master and ninj) else None
        orc.step(n, u)
        out = ks.step(n, u)
        assert out.status == 0
        assert out.n_alive == orc.count(), f"step {n}: alive {out.n_alive} vs {orc.count()}"
        if not (master or n == 1):
            continue
        ma, mb = orc.master(), ks.master(0)
        dc = np.abs(ma["nq_counts"] - mb["nq_counts"]).sum()
        dph = np.abs(ma["nph_counts"] - mb["nph_counts"]).sum()
        assert dc <= max_dcount and dph <= 2 * max_dcount, f"step {n}: count differences {dc}, {dph}"
        assert np.array_equal(ma["esc"][:2], mb["esc"][:2]), f"step {n}: escapes differ"
        for k in MOMENTS:
            e = _l2rel(ma[k], mb[k])
            worst[k] = max(worst[k], e)
            assert e <= tol_l2, f"step {n}: {k} relative L2 error {e:.2e}"
    a, b = orc.particles(), ks.particles(0)
    assert len(a["x"]) == len(b["x"])
    # scales over all three components (at phi = pi/2, x and Vx are rounding noise)
    for keys, tol in ((("x", "y", "z"), 1e-9), (("Vx", "Vy", "Vz"), 1e-6)):
        sc = max(np.abs(a[k]).max() for k in keys)
        for k in keys:
            assert np.abs(a[k] - b[k]).max() / sc < tol, k
    assert np.mean(a["qind"] == b["qind"]) > 0.999
    assert np.mean(a["vparind"] == b["vparind"]) > 0.999
    assert np.array_equal(a["vperpind"], b["vperpind"])
    ks.close()
    orc.close()
    return worst


def test_mixed_vs_oracle_config1_every_step_master():
    worst = _run(grids.config1(n_total=20000), 12)
    print("worst relative L2 errors:", {k: f"{v:.1e}" for k, v in worst.items()})


def test_mixed_vs_oracle_ndatfac3_ub_injection_uniform_stream():
    tube = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=15000, ratio=1.08)
    tube.n_inj_ub = 17
    tube.n_inj_lb = 23
    _run(tube, 13, ndatfac=3, uniforms=True)


def test_mixed_vs_oracle_south_hemisphere_eambsign0():
    tube = grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=8000, south=True, p=6.0)
    tube.eambsign = 0
    tube.n_inj_ub = 5
    _run(tube, 9, ndatfac=2)


def test_mixed_vs_oracle_full_grid_1e6_ions():
    # the 256-cell, 64x64 grid of the full-size workload with 1e6 ions
    worst = _run(grids.config2(n_total=1_000_000), 10, max_dcount=6, tol_l2=1e-5)
    print("worst relative L2 errors:", {k: f"{v:.1e}" for k, v in worst.items()})


def test_mixed_vs_oracle_with_reorder_every_master_step():
    _run(grids.config1(n_total=20000), 12, reorder_every=1)
