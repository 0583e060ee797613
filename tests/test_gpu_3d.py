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
