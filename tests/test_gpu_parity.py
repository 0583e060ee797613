"""FP64 parity of the HIP path against the oracle, through the C-ABI.

Bar: bit-exact.  The oracle and the kernels evaluate the same IEEE-754
operation sequence for the push, the field and the moments (shared
kaos_dipole.h, contraction off on both compilers), and integer counts are
exact, so every per-ion double, every count, every moment and every field
value must be identical after every step.  Independent parts (cell scans,
removal/packing, injection slots, histogram, field lag, master cadence) are
written separately on the two sides.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
MASTER_KEYS = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp",
               "M2par", "EA", "EG", "esc")


def _run_pair(tube, nsteps, ndatfac=1, seed=11, ions_seed=3, uniforms=False, cap_extra=20000,
              reorder_every=0):
    q0, v0, w0 = grids.initial_ions(tube, seed=ions_seed)
    cap = len(q0) + cap_extra
    orc = oracle.Oracle(tube, q0, v0, w0, ndatfac=ndatfac, seed=seed, capacity=cap)
    ks = KineticSolver([tube], precision=KAOS_FP64, ndatfac=ndatfac, seed=seed, capacity=[cap],
                       reorder_every=reorder_every)
    ks.upload(0, q0, v0, w0)
    rng = np.random.default_rng(99)
    ninj = (tube.n_inj_lb if tube.lbcond else 0) + (tube.n_inj_ub if tube.ubcond else 0)
    for n in range(1, nsteps + 1):
        master = n >= 2 and n % ndatfac == 0
        u = None
        if uniforms and master and ninj:
            u = 1.0 - rng.random(6 * ninj)          # (0, 1]
        orc.step(n, u)
        out = ks.step(n, u)
        assert out.status == 0
        yield n, master or n == 1, orc, ks


def _compare(orc, ks, n, check_master):
    a, b = orc.particles(), ks.particles(0)
    assert len(a["x"]) == len(b["x"]), f"step {n}: ion count {len(a['x'])} vs {len(b['x'])}"
    for k in PART_KEYS:
        if not np.array_equal(a[k], b[k]):
            bad = np.nonzero(a[k] != b[k])[0]
            raise AssertionError(f"step {n}: {k} differs at {len(bad)} ions, first {bad[:5]}: "
                                 f"{a[k][bad[:3]]} vs {b[k][bad[:3]]}")
    if check_master:
        ma, mb = orc.master(), ks.master(0)
        for k in MASTER_KEYS:
            assert np.array_equal(ma[k], mb[k]), f"step {n}: master output {k} differs"


def test_fp64_bitexact_config1_every_step_master():
    tube = grids.config1(n_total=20000)
    for n, master, orc, ks in _run_pair(tube, 12):
        _compare(orc, ks, n, master)


def test_fp64_bitexact_ndatfac3_ub_injection_uniform_stream():
    tube = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=15000, ratio=1.08)
    tube.n_inj_ub = 17
    tube.n_inj_lb = 23
    for n, master, orc, ks in _run_pair(tube, 13, ndatfac=3, uniforms=True):
        _compare(orc, ks, n, master)


def test_fp64_bitexact_south_hemisphere_eambsign0():
    tube = grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=8000, south=True, p=6.0)
    tube.eambsign = 0
    tube.n_inj_ub = 5
    for n, master, orc, ks in _run_pair(tube, 9, ndatfac=2):
        _compare(orc, ks, n, master)


def test_fp64_bitexact_with_reorder_every_master_step():
    # compaction + sort by cell after every master step: slot order changes,
    # results and the downloaded (id-ordered) ions must not
    tube = grids.config1(n_total=20000)
    for n, master, orc, ks in _run_pair(tube, 12, reorder_every=1):
        _compare(orc, ks, n, master)


def test_fp64_bitexact_with_reorder_ndatfac3_uniform_stream():
    tube = grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=15000, ratio=1.08)
    tube.n_inj_ub = 17
    tube.n_inj_lb = 23
    for n, master, orc, ks in _run_pair(tube, 13, ndatfac=3, uniforms=True, reorder_every=1):
        _compare(orc, ks, n, master)
