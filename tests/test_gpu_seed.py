"""Initial ions generated on the device (kaos_seed_particles) against the
oracle's restatement of DensityProfileBSub + VelocityDistributionSub +
InitialConditionsSub (orc_seed_ions), through the C-ABI.

FP64: the seeded state and every later step are bit-exact against the oracle
started from the oracle's own initial ions.  Mixed: seeding equals uploading
the same ions (kaos_upload_particles + kaos_upload_vperp) bit for bit, in the
state and in the moments after the steps that follow.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
# before the first step the oracle has no cell indices yet (KSA computes the
# k0 cells at step 1); the library holds them from the start
STATE_KEYS = PART_KEYS[:9]
MASTER_KEYS = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp",
               "M2par", "EA", "EG", "esc")


def _counts_temp(t, zero=(), tmin=2000.0, tmax=6000.0):
    counts = t.cell_counts.copy()
    for Q in zero:
        counts[Q] = 0
    return counts, np.linspace(tmin, tmax, t.nq)


def _same(a, b, keys, what):
    for k in keys:
        if not np.array_equal(a[k], b[k]):
            bad = np.nonzero(np.asarray(a[k]) != np.asarray(b[k]))
            raise AssertionError(f"{what}: {k} differs at {len(bad[0])} entries")


def _cap(t, counts, nsteps):
    return int(counts.sum()) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 512


def _oracles(tubes, seeds, nsteps, seed, ndatfac=1):
    orcs = []
    for i, (t, (counts, temp)) in enumerate(zip(tubes, seeds)):
        s = oracle.seed_ions(t, counts, temp, seed, tube_index=i)
        o = oracle.Oracle(t, s["q0"], s["vpar"], s["vperp"], ndatfac=ndatfac, seed=seed,
                          capacity=_cap(t, counts, nsteps), tube_index=i)
        o.set_vperp(s["v1"], s["v2"])
        orcs.append(o)
    return orcs


def test_seed_fp64_bitexact_then_steps():
    t = grids.dipole_tube(nq=64, nvpar=40, nvperp=40, n_total=30000, ratio=1.05)
    counts, temp = _counts_temp(t, zero=(7, 8, 62, 63))
    nsteps, seed = 10, 9001
    print("seed fp64: tube 64 cells, ions", int(counts.sum()), flush=True)
    orcs = _oracles([t], [(counts, temp)], nsteps, seed)
    ks = KineticSolver([t], precision=KAOS_FP64, seed=seed, capacity=[_cap(t, counts, nsteps)])
    ks.seed(0, counts, temp, seed)
    assert ks.count(0) == counts.sum()
    P = ks.particles(0)
    _same(orcs[0].particles(), P, STATE_KEYS, "seeded state")
    assert np.array_equal(P["qind"], np.repeat(np.arange(t.nq), counts) + 1)
    for n in range(1, nsteps + 1):
        orcs[0].step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == orcs[0].count()
        _same(orcs[0].particles(), ks.particles(0), PART_KEYS, f"step {n}")
        if n >= 2:
            _same(orcs[0].master(), ks.master(0), MASTER_KEYS, f"master {n}")
    ks.close()


def test_seed_fp64_batched_tubes_bitexact():
    tubes = [grids.config1(n_total=8000),
             grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=5000, south=True, p=6.0),
             grids.dipole_tube(nq=48, nvpar=40, nvperp=40, n_total=6000, p=5.0, r_ub=3.0)]
    seeds = [_counts_temp(tubes[0]), _counts_temp(tubes[1], zero=(0, 31), tmin=800.0, tmax=900.0),
             _counts_temp(tubes[2], zero=tuple(range(20, 48)), tmin=9000.0, tmax=3000.0)]
    nsteps, seed = 6, 77
    print("seed fp64 batched: 3 tubes", [int(c.sum()) for c, _ in seeds], flush=True)
    orcs = _oracles(tubes, seeds, nsteps, seed)
    ks = KineticSolver(tubes, precision=KAOS_FP64, seed=seed,
                       capacity=[_cap(t, c, nsteps) for t, (c, _) in zip(tubes, seeds)])
    for i, (c, T) in enumerate(seeds):
        ks.seed(i, c, T, seed)
    for i, o in enumerate(orcs):
        _same(o.particles(), ks.particles(i), STATE_KEYS, f"seeded tube {i}")
    for n in range(1, nsteps + 1):
        for o in orcs:
            o.step(n)
        assert ks.step(n).status == 0
        if n >= 2:
            for i, o in enumerate(orcs):
                _same(o.master(), ks.master(i), MASTER_KEYS, f"master {n} tube {i}")
                _same(o.particles(), ks.particles(i), PART_KEYS, f"step {n} tube {i}")
    ks.close()


def test_seed_mixed_equals_upload():
    t = grids.config1(n_total=40000)
    counts, temp = _counts_temp(t, zero=(30,))
    nsteps, seed = 8, 5150
    print("seed mixed vs upload: ions", int(counts.sum()), flush=True)
    s = oracle.seed_ions(t, counts, temp, seed)
    cap = [_cap(t, counts, nsteps)]
    ka = KineticSolver([t], precision=KAOS_MIXED, seed=seed, capacity=cap)
    kb = KineticSolver([t], precision=KAOS_MIXED, seed=seed, capacity=cap)
    ka.seed(0, counts, temp, seed)
    kb.upload(0, s["q0"], s["vpar"], s["vperp"])
    kb.upload_vperp(0, s["v1"], s["v2"])
    _same(ka.particles(0), kb.particles(0), PART_KEYS, "seeded vs uploaded")
    for n in range(1, nsteps + 1):
        assert ka.step(n).status == 0 and kb.step(n).status == 0
        if n >= 2:
            _same(ka.master(0), kb.master(0), MASTER_KEYS, f"master {n}")
    _same(ka.particles(0), kb.particles(0), PART_KEYS, "after the steps")
    ka.close()
    kb.close()


def test_seed_edge_cases_and_errors():
    t = grids.dipole_tube(nq=24, nvpar=16, nvperp=16, n_total=2000, ratio=1.1)
    t.n_inj_lb, t.n_inj_ub = 13, 4
    counts, temp = _counts_temp(t)
    ks = KineticSolver([t], precision=KAOS_FP64, seed=3, capacity=[int(counts.sum()) + 200])
    bad = counts.copy()
    bad[4] = -1
    with pytest.raises(KaosError):
        ks.seed(0, bad, temp, 1)
    big = counts.copy()
    big[0] += 10 ** 6
    with pytest.raises(KaosError):
        ks.seed(0, big, temp, 1)
    cold = temp.copy()
    cold[2] = 0.0
    with pytest.raises(KaosError):
        ks.seed(0, counts, cold, 1)
    # no ions at all: the tube fills by injection alone
    ks.seed(0, np.zeros(t.nq, np.int64), np.zeros(t.nq), 1)
    assert ks.count(0) == 0
    o = oracle.Oracle(t, np.zeros(0), np.zeros(0), np.zeros(0), seed=3, capacity=int(counts.sum()) + 200)
    for n in range(1, 5):
        o.step(n)
        assert ks.step(n).status == 0
    assert ks.count(0) == o.count() > 0
    _same(o.particles(), ks.particles(0), PART_KEYS, "injection only")
    with pytest.raises(KaosError):
        ks.seed(0, counts, temp, 1)             # after the first step
    ks.close()
