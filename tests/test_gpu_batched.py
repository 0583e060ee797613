"""Many flux tubes in one context (one launch per kernel per step): uneven ion
counts, grid sizes, L-shells and hemispheres.  FP64 is bit-exact against one
oracle per tube (tubes are independent; the injection stream is keyed by the
tube index); the mixed form stays within the single-tube bounds."""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
MASTER_KEYS = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp",
               "M2par", "EA", "EG", "esc")


def _tubes():
    t = [grids.config1(n_total=9000),
         # > 64 Vpar cells: moments by the block-per-cell kernel (row groups G = 2)
         grids.dipole_tube(nq=100, nvpar=100, nvperp=24, n_total=15000, p=5.0, r_ub=3.5, ratio=1.02, T_eV=0.6),
         grids.dipole_tube(nq=24, nvpar=16, nvperp=12, n_total=700, ratio=1.1),          # smaller than one tile
         grids.dipole_tube(nq=48, nvpar=40, nvperp=40, n_total=12000, south=True, p=6.0, T_eV=0.2),
         grids.dipole_tube(nq=64, nvpar=64, nvperp=64, n_total=5000, p=3.5, r_ub=2.6, ratio=1.03)]
    t[1].n_inj_ub = 11
    t[3].eambsign = 0
    t[4].lbcond = 0
    t[4].n_inj_lb = 0
    return t


def _run(precision, nsteps=10, ndatfac=1, reorder_every=0, seed=17):
    tubes = _tubes()
    ions = [grids.initial_ions(t, seed=40 + i) for i, t in enumerate(tubes)]
    caps = [len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 512 for t, (q, _, _) in zip(tubes, ions)]
    orcs = [oracle.Oracle(t, *ions[i], ndatfac=ndatfac, seed=seed, capacity=caps[i], tube_index=i)
            for i, t in enumerate(tubes)]
    ks = KineticSolver(tubes, precision=precision, ndatfac=ndatfac, seed=seed, capacity=caps,
                       reorder_every=reorder_every)
    for i, (q, v, w) in enumerate(ions):
        ks.upload(i, q, v, w)
    for n in range(1, nsteps + 1):
        for o in orcs:
            o.step(n)
        out = ks.step(n)
        assert out.status == 0
        assert out.n_alive == sum(o.count() for o in orcs)
        yield n, n == 1 or (n >= 2 and n % ndatfac == 0), tubes, orcs, ks
    ks.close()
    for o in orcs:
        o.close()


@pytest.mark.parametrize("ndatfac,reorder", [(1, 0), (3, 1)])
def test_batched_fp64_bitexact_per_tube(ndatfac, reorder):
    for n, master, tubes, orcs, ks in _run(KAOS_FP64, ndatfac=ndatfac, reorder_every=reorder):
        for i, o in enumerate(orcs):
            a, b = o.particles(), ks.particles(i)
            assert len(a["x"]) == len(b["x"]), (n, i)
            for k in PART_KEYS:
                assert np.array_equal(a[k], b[k]), f"step {n} tube {i}: {k}"
            if master:
                ma, mb = o.master(), ks.master(i)
                for k in MASTER_KEYS:
                    assert np.array_equal(ma[k], mb[k]), f"step {n} tube {i}: {k}"


def test_batched_mixed_close_to_oracle():
    for n, master, tubes, orcs, ks in _run(KAOS_MIXED):
        if not master:
            continue
        for i, o in enumerate(orcs):
            ma, mb = o.master(), ks.master(i)
            assert np.abs(ma["nq_counts"] - mb["nq_counts"]).sum() <= 2, (n, i)
            for k in ("M0", "M1par", "M2", "EA"):
                na = np.linalg.norm(ma[k])
                assert np.linalg.norm(ma[k] - mb[k]) <= 1e-4 * max(na, 1e-300), (n, i, k)
