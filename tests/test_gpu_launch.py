"""Launch shape, autotuning and graph replay do not change results: FP64 stays
bit-exact against the oracle (a) for forced launch shapes, grid and
persistent, with and without the LDS histogram window, (b) when kaos_run
replays a captured hipGraph per step, checked after every replay with eager
downloads in between, and (c) across the online autotuner's candidate steps
(tiles per group, LDS window, grid or persistent launch, 256 or 512 threads
per block); the mixed push gives the same results with 256 and 512 threads."""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64, KAOS_MIXED, KAOS_FLAG_NO_AUTOTUNE, KAOS_FLAG_NO_GRAPH

pytestmark = pytest.mark.gpu

MASTER_KEYS = ("nq_counts", "nph_counts", "M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA", "esc")
PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "qind", "vparind", "vperpind")


def _pair(tubes, flags, ndatfac=1, seed=5, reorder_every=0):
    ions = [grids.initial_ions(t, seed=60 + i) for i, t in enumerate(tubes)]
    caps = [len(q) + 64 * (t.n_inj_lb + t.n_inj_ub) + 512 for t, (q, _, _) in zip(tubes, ions)]
    orcs = [oracle.Oracle(t, *ions[i], ndatfac=ndatfac, seed=seed, capacity=caps[i], tube_index=i)
            for i, t in enumerate(tubes)]
    ks = KineticSolver(tubes, precision=KAOS_FP64, ndatfac=ndatfac, seed=seed, capacity=caps, flags=flags,
                       reorder_every=reorder_every)
    for i, (q, v, w) in enumerate(ions):
        ks.upload(i, q, v, w)
    return orcs, ks


def _check(orcs, ks, n, master=True):
    for i, o in enumerate(orcs):
        a, b = o.particles(), ks.particles(i)
        for k in PART_KEYS:
            assert np.array_equal(a[k], b[k]), f"step {n} tube {i}: {k}"
        if master:
            ma, mb = o.master(), ks.master(i)
            for k in MASTER_KEYS:
                assert np.array_equal(ma[k], mb[k]), f"step {n} tube {i}: {k}"


@pytest.mark.parametrize("shape,block", [((1, 0, 0), 256), ((4, 2048, 1), 512), ((32, 6144, 1), 256),
                                         ((63, 12288, 0), 512), ((8, 4096, 1), 512)])
def test_forced_launch_shapes_bitexact(shape, block):
    tubes = [grids.config1(n_total=30000), grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=9000, ratio=1.08)]
    orcs, ks = _pair(tubes, KAOS_FLAG_NO_AUTOTUNE)
    ks.set_launch(*shape)
    ks.set_launch_block(block)
    assert ks.get_launch()[:3] == shape and ks.get_launch_block() == block
    for n in range(1, 7):
        for o in orcs:
            o.step(n)
        assert ks.step(n).status == 0
        _check(orcs, ks, n)
    ks.close()


@pytest.mark.parametrize("ndatfac", [1, 2])
def test_graph_replay_bitexact_every_step(ndatfac):
    tubes = [grids.config1(n_total=20000), grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=6000,
                                                           south=True, p=6.0)]
    orcs, ks = _pair(tubes, KAOS_FLAG_NO_AUTOTUNE, ndatfac=ndatfac, reorder_every=3)
    for o in orcs:
        o.step(1)
    ks.step(1)
    # kaos_run of one step at a time replays the captured graph of that step kind
    for n in range(2, 16):
        for o in orcs:
            o.step(n)
        ks.run(n, 1)
        _check(orcs, ks, n, master=(n % ndatfac == 0))
    ks.close()


def test_autotuned_run_bitexact_and_settles():
    tubes = [grids.config1(n_total=40000)]
    orcs, ks = _pair(tubes, 0)
    for o in orcs:
        o.step(1)
    ks.step(1)
    assert ks.get_launch()[3] == 1          # still tuning
    n = 2
    while n < 56:                           # 2 warm-up + 3 x 16 candidate steps + the choice + graph steps
        for o in orcs:
            o.step(n)
        ks.run(n, 1)
        _check(orcs, ks, n)
        n += 1
    tpb, hw, pers, tuning = ks.get_launch()
    assert tuning == 0 and tpb in (8, 16) and hw in (4096, 6144) and pers in (0, 1)
    assert ks.get_launch_block() in (256, 512)
    ks.close()


def test_mixed_graph_and_tuning_match_eager_untuned():
    # the mixed form through graphs + autotuning equals itself run eagerly untuned
    tube = grids.config1(n_total=30000)
    q0, v0, w0 = grids.initial_ions(tube, seed=3)
    cap = len(q0) + 40 * tube.n_inj_lb + 512
    a = KineticSolver([tube], precision=KAOS_MIXED, seed=2, capacity=[cap])
    b = KineticSolver([tube], precision=KAOS_MIXED, seed=2, capacity=[cap],
                      flags=KAOS_FLAG_NO_GRAPH | KAOS_FLAG_NO_AUTOTUNE)
    for s in (a, b):
        s.upload(0, q0, v0, w0)
        s.step(1)
    for n in range(2, 28):
        a.run(n, 1)
        b.step(n)
        ma, mb = a.master(0), b.master(0)
        for k in MASTER_KEYS:
            assert np.array_equal(ma[k], mb[k]), f"step {n}: {k}"
    pa, pb = a.particles(0), b.particles(0)
    for k in PART_KEYS:
        assert np.array_equal(pa[k], pb[k]), k
    a.close()
    b.close()


def test_mixed_push_block_sizes_agree():
    """The mixed push with 512 threads per block (two ions per thread) gives
    the same state and outputs as with 256 (four per thread)."""
    t = grids.config1(n_total=40000)
    q, v, w = grids.initial_ions(t, seed=9)
    cap = len(q) + 20 * (t.n_inj_lb + t.n_inj_ub) + 512
    runs = []
    for block in (256, 512):
        ks = KineticSolver([t], precision=KAOS_MIXED, ndatfac=2, seed=3, capacity=[cap], flags=KAOS_FLAG_NO_AUTOTUNE)
        ks.upload(0, q, v, w)
        ks.set_launch_block(block)
        for n in range(1, 11):
            assert ks.step(n).status == 0
        runs.append((ks.particles(0), ks.master(0)))
        ks.close()
    (pa, ma), (pb, mb) = runs
    for k in PART_KEYS:
        assert np.array_equal(pa[k], pb[k]), k
    for k in MASTER_KEYS:
        assert np.array_equal(ma[k], mb[k]), k
