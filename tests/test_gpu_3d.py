"""The 3-D velocity grid (ION2VPERPflag = 1; KAOS_FLAG_VPERP3D) on the HIP path
against the oracle, through the C-ABI.

FP64: bit-exact after every step: the particles, and at master steps the 3-D
counts [Q][V1][V2][Vpar], the eight 3-D moments, the moment block the next
master step's fields read (M0, M1Par, M2, M2Par; M1Perp and M2Perp zero), the
config counts and the fields; for two tubes of one context (one southern)
with a drift in Vperp2, ndatfac 2 and the seeded injection stream; with ICR
heating on (the counts take the pre-kick components) and a reorder every
master step.  Mixed: config counts within a few ions, the 3-D counts within a
few dozen, the moments within 1e-6 (relative L2).  Errors: the 3-D download
from a 2-D context, a context mixing 2-D and 3-D tubes, a missing 3-D array.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
MASTER_KEYS = ("nq_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA", "EG", "esc")


def _tubes3():
    a = grids.with_vgrid3(grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=12000, ratio=1.08), 14, 12,
                          vperp_max=6.0e3)
    a.n_inj_ub = 7
    b = grids.with_vgrid3(grids.dipole_tube(nq=32, nvpar=80, nvperp=16, n_total=6000, south=True, p=6.0), 10, 9,
                          vperp_max=9.0e3)
    return [a, b]


def _pair(tubes, precision, ndatfac=2, seed=13, reorder_every=0, nsteps=10):
    ions = [grids.initial_ions(t, seed=70 + i) for i, t in enumerate(tubes)]
    caps = [len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 512 for t, (q, _, _) in zip(tubes, ions)]
    orcs = [oracle.Oracle(t, *ions[i], ndatfac=ndatfac, seed=seed, capacity=caps[i], tube_index=i)
            for i, t in enumerate(tubes)]
    ks = KineticSolver(tubes, precision=precision, ndatfac=ndatfac, seed=seed, capacity=caps,
                       reorder_every=reorder_every)
    rng = np.random.default_rng(9)
    for i, (q, v, w) in enumerate(ions):
        ks.upload(i, q, v, w)
        ph = rng.uniform(0.0, 2.0 * np.pi, len(w))
        v1, v2 = w * np.cos(ph), w * np.sin(ph) + 250.0      # a drift in Vperp2
        ks.upload_vperp(i, v1, v2)
        orcs[i].set_vperp(v1, v2)
    return orcs, ks


def _steps(orcs, ks, nsteps, ndatfac):
    for n in range(1, nsteps + 1):
        for o in orcs:
            o.step(n)
        out = ks.step(n)
        assert out.status == 0
        assert out.n_alive == sum(o.count() for o in orcs)
        yield n, n == 1 or (n >= 2 and n % ndatfac == 0)


def _bitexact(orcs, ks, n, master):
    for i, o in enumerate(orcs):
        a, b = o.particles(), ks.particles(i)
        assert len(a["x"]) == len(b["x"]), (n, i)
        for k in PART_KEYS:
            if not np.array_equal(a[k], b[k]):
                bad = np.nonzero(a[k] != b[k])[0]
                raise AssertionError(f"step {n} tube {i}: {k} differs at {len(bad)} ions")
        if master:
            ma, mb = o.master(), ks.master(i)
            for k in MASTER_KEYS:
                assert np.array_equal(ma[k], mb[k]), f"step {n} tube {i}: {k}"
            assert not mb["M1perp"].any() and not mb["M2perp"].any()
            a3, b3 = o.master3(), ks.master3(i)
            assert np.array_equal(a3["n3_counts"], b3["n3_counts"]), f"step {n} tube {i}: 3-D counts"
            for k in oracle.MOMENTS3:
                assert np.array_equal(a3[k], b3[k]), f"step {n} tube {i}: {k}"


def test_3d_fp64_bitexact_two_tubes_ndatfac2():
    tubes = _tubes3()
    orcs, ks = _pair(tubes, KAOS_FP64)
    print("3-D fp64: 2 tubes (14x12x24, 10x9x80), ndatfac 2, 10 steps", flush=True)
    nmaster = 0
    for n, master in _steps(orcs, ks, 10, 2):
        _bitexact(orcs, ks, n, master)
        nmaster += master
    assert nmaster == 6
    m3 = ks.master3(0)
    assert m3["n3_counts"].sum() > 0 and m3["M2perp2"].any()
    # the drift in Vperp2 shows in M1Perp2 (density-weighted mean over the tube)
    assert abs(np.average(m3["M1perp2"], weights=m3["M0"]) - 250.0) < 100.0
    ks.close()


def test_3d_fp64_bitexact_heating_and_reorder():
    t = grids.with_vgrid3(grids.config1(n_total=15000), 16, 16, vperp_max=5.0e3)
    t.icr = 1
    orcs, ks = _pair([t], KAOS_FP64, ndatfac=1, reorder_every=1, nsteps=8)
    print("3-D fp64: config-1 tube with ICR heating, reorder every master step, 8 steps", flush=True)
    for n, master in _steps(orcs, ks, 8, 1):
        _bitexact(orcs, ks, n, master)
    ks.close()


def test_3d_mixed_close_to_oracle():
    tubes = _tubes3()
    orcs, ks = _pair(tubes, KAOS_MIXED, ndatfac=1, nsteps=8)
    print("3-D mixed: 2 tubes, 8 steps", flush=True)
    for n, master in _steps(orcs, ks, 8, 1):
        for i, o in enumerate(orcs):
            ma, mb = o.master(), ks.master(i)
            a3, b3 = o.master3(), ks.master3(i)
            assert np.abs(ma["nq_counts"] - mb["nq_counts"]).sum() <= 4, (n, i)
            assert np.abs(a3["n3_counts"] - b3["n3_counts"]).sum() <= 40, (n, i)
            for k in oracle.MOMENTS3:
                na = np.linalg.norm(a3[k])
                assert np.linalg.norm(a3[k] - b3[k]) <= 1e-6 * na, (n, i, k)
    ks.close()


def test_3d_errors():
    t2 = grids.dipole_tube(nq=24, nvpar=16, nvperp=16, n_total=2000, ratio=1.1)
    q, v, w = grids.initial_ions(t2, seed=1)
    ks = KineticSolver([t2], precision=KAOS_FP64, ndatfac=1, seed=1, capacity=[len(q) + 512])
    ks.upload(0, q, v, w)
    assert ks.step(1).status == 0
    with pytest.raises(KaosError) as e:            # a 2-D context has no 3-D outputs
        ks.master3(0)
    assert e.value.code == 4
    ks.close()
    t3 = grids.with_vgrid3(grids.dipole_tube(nq=24, nvpar=16, nvperp=16, n_total=2000, ratio=1.1), 8, 8,
                           vperp_max=5.0e3)
    with pytest.raises(ValueError):                # 2-D and 3-D tubes in one context
        KineticSolver([t3, t2], precision=KAOS_FP64, ndatfac=1, seed=1, capacity=[4096, 4096])
    t3.arrays["d3v3"] = None
    with pytest.raises(KaosError) as e:             # a missing 3-D array
        KineticSolver([t3], precision=KAOS_FP64, ndatfac=1, seed=1, capacity=[4096])
    assert e.value.code == 1


# ---------------------------------------------------------------------------
# Full size on the 3-D grid: the benchmark tube (256 cells, 64 Vpar cells) with
# 32 x 32 (Vperp1, Vperp2) cells and 1e8 ions seeded on the device, where the
# oracle would take hours, checked through properties that hold at any size.

N3_FULL = 100_000_000
N3_STEPS = 6


def _full3(precision):
    tube = grids.with_vgrid3(grids.config2(n_total=N3_FULL), 32, 32)
    counts = tube.cell_counts
    cap = int(counts.sum()) + (N3_STEPS + 2) * (tube.n_inj_lb + tube.n_inj_ub) + 1024
    ks = KineticSolver([tube], precision=precision, ndatfac=1, seed=5, capacity=[cap])
    ks.seed(0, counts, np.full(tube.nq, tube.Ts_lb), 20241)
    steps = []
    for n in range(1, N3_STEPS + 1):
        out = ks.step(n)
        assert out.status == 0
        m = ks.master(0)
        m3 = ks.master3(0)
        steps.append((out.n_alive, out.n_injected, m, m3))
    ks.close()
    return tube, steps


@pytest.fixture(scope="module")
def full3():
    print(f"3-D full size: {N3_FULL} ions, 32x32x64 cells, {N3_STEPS} steps, fp64 then mixed", flush=True)
    return {"fp64": _full3(KAOS_FP64), "mixed": _full3(KAOS_MIXED)}


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_3d_full_size_counts_and_bookkeeping(full3, form):
    tube, st = full3[form]
    k = tube.nsnormfac / tube.arrays["d3xC"]
    # step 1: every seeded ion is in its cell; the 3-D counts hold those inside the grid
    m, m3 = st[0][2], st[0][3]
    assert np.array_equal(m["nq_counts"], tube.cell_counts)
    inside = m3["n3_counts"].reshape(tube.nq, -1).sum(axis=1)
    assert np.all(inside <= m["nq_counts"]) and inside.sum() >= 0.999 * N3_FULL
    removed = [int(s[2]["esc"][0] + s[2]["esc"][1]) for s in st]
    for i in range(1, N3_STEPS - 1):
        alive, inj, m, m3 = st[i]
        assert alive == st[i - 1][0] - removed[i] + inj, i + 1
        assert m["nq_counts"].sum() == alive - removed[i + 1], i + 1
    for _, _, m, m3 in st[1:]:
        assert np.array_equal(m["NqRT"], m["nq_counts"] * k)
        inside = m3["n3_counts"].reshape(tube.nq, -1).sum(axis=1)
        assert np.allclose(m3["M0"], inside * k, rtol=1e-12, atol=0)
        assert np.array_equal(m["M0"], m3["M0"]) and np.array_equal(m["M2par"], m3["M2par"])


def test_3d_full_size_mixed_within_1e6_of_fp64(full3):
    tube = full3["fp64"][0]
    a, b = full3["fp64"][1][-1][3], full3["mixed"][1][-1][3]
    for key in oracle.MOMENTS3:
        # M1Perp1 and M1Perp2 are means of signed components, ~0 up to sampling
        # noise for this isotropic population: their error is measured against
        # the perpendicular thermal speed sqrt(2 M2Perp / m).  The mixed form
        # bins Vperp1/Vperp2 from cells found in binary64 (unheated: they never
        # change), so what moves these means is an ion whose config or Vpar
        # cell differs after the binary32 stages: it takes its whole V1 to the
        # neighbouring config cell, a shift of ~sigma/count against a mean
        # that is itself only noise
        if key in ("M1perp1", "M1perp2"):
            scale = np.linalg.norm(np.sqrt(2.0 * a["M2perp" + key[-1]] / tube.ms))
        else:
            scale = np.linalg.norm(a[key])
        err = np.linalg.norm(b[key] - a[key]) / scale
        assert err <= 1e-6, (key, err)
    assert np.abs(full3["fp64"][1][-1][2]["nq_counts"] - full3["mixed"][1][-1][2]["nq_counts"]).sum() <= 16
    # the (V1, V2) marginal summed over config and Vpar cells: the same binary64
    # perpendicular cells in both forms, so it differs only by the few ions
    # inside the config or Vpar grid in one form and outside it in the other
    ma = a["n3_counts"].sum(axis=(0, 3))
    mb = b["n3_counts"].sum(axis=(0, 3))
    assert np.abs(ma - mb).sum() <= 16, np.abs(ma - mb).sum()


def test_3d_fp64_bitexact_grid_larger_than_the_window():
    # 40 x 40 x 48 = 76800 bins per config cell: more than the binning pass's
    # LDS window holds, so every count goes through global atomics
    t = grids.with_vgrid3(grids.dipole_tube(nq=24, nvpar=48, nvperp=16, n_total=10000, ratio=1.1), 40, 40,
                          vperp_max=7.0e3)
    orcs, ks = _pair([t], KAOS_FP64, ndatfac=1, nsteps=5)
    print("3-D fp64: 40x40x48 grid (no LDS window), 5 steps", flush=True)
    for n, master in _steps(orcs, ks, 5, 1):
        _bitexact(orcs, ks, n, master)
    assert ks.master3(0)["n3_counts"].sum() > 0
    ks.close()
