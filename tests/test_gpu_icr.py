"""ICR wave heating (WaveHeatingSub, _8_1_1_0_WaveHeating.f90:32-314) on the
HIP path against the oracle, through the C-ABI.

FP64: bit-exact after every step, as the rest of the path (the kick is the
shared kaos_icr_kick on both sides, with the same Philox4x32 stream keyed by
seed, step, tube and ion id).  Mixed: the kick is evaluated in binary64 from
binary32 components, so the bar is closeness (stated below).  The kicks are
~10 m/s per step at the reference's heating parameters on the config-1 tube,
so Vperp cells change within the 12 steps run here.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART_KEYS = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
MASTER_KEYS = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp",
               "M2par", "EA", "EG", "esc")


def _heated(tube, **kw):
    tube.icr = 1
    for k, v in kw.items():
        setattr(tube, k, v)
    return tube


def _pair(tubes, precision, ndatfac=1, seed=21, reorder_every=0, nsteps=12, components=False):
    ions = [grids.initial_ions(t, seed=60 + i) for i, t in enumerate(tubes)]
    caps = [len(q) + (nsteps + 2) * (t.n_inj_lb + t.n_inj_ub) + 512 for t, (q, _, _) in zip(tubes, ions)]
    orcs = [oracle.Oracle(t, *ions[i], ndatfac=ndatfac, seed=seed, capacity=caps[i], tube_index=i)
            for i, t in enumerate(tubes)]
    ks = KineticSolver(tubes, precision=precision, ndatfac=ndatfac, seed=seed, capacity=caps,
                       reorder_every=reorder_every)
    rng = np.random.default_rng(5)
    for i, (q, v, w) in enumerate(ions):
        ks.upload(i, q, v, w)
        if components:
            ph = rng.uniform(0.0, 2.0 * np.pi, len(w))
            v1, v2 = w * np.cos(ph), w * np.sin(ph)
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
                raise AssertionError(f"step {n} tube {i}: {k} differs at {len(bad)} ions: "
                                     f"{a[k][bad[:3]]} vs {b[k][bad[:3]]}")
        if master:
            ma, mb = o.master(), ks.master(i)
            for k in MASTER_KEYS:
                assert np.array_equal(ma[k], mb[k]), f"step {n} tube {i}: {k}"


def test_icr_fp64_bitexact_config1_every_step():
    tube = _heated(grids.config1(n_total=20000))
    orcs, ks = _pair([tube], KAOS_FP64)
    w0 = orcs[0].particles()["vperp"].copy()
    for n, master in _steps(orcs, ks, 12, 1):
        _bitexact(orcs, ks, n, master)
    # the heating acted: Vperp moved for most ions and some changed Vperp cell
    a = orcs[0].particles()
    k = min(len(w0), len(a["vperp"]))
    assert np.mean(a["vperp"][:k] != w0[:k]) > 0.9
    ks.close()


def test_icr_fp64_bitexact_components_ndatfac2_reorder():
    tube = _heated(grids.dipole_tube(nq=40, nvpar=24, nvperp=20, n_total=12000, ratio=1.08),
                   icr_lambda_perp=2000.0, icr_chi_perp2=2.0)
    tube.n_inj_ub = 9
    orcs, ks = _pair([tube], KAOS_FP64, ndatfac=2, reorder_every=1, components=True)
    for n, master in _steps(orcs, ks, 12, 2):
        _bitexact(orcs, ks, n, master)
    ks.close()


def test_icr_fp64_bitexact_batched_mixed_heating():
    tubes = [_heated(grids.config1(n_total=6000)),
             grids.dipole_tube(nq=32, nvpar=16, nvperp=16, n_total=4000, south=True, p=6.0),
             _heated(grids.dipole_tube(nq=48, nvpar=40, nvperp=40, n_total=5000, p=5.0, r_ub=3.0),
                     icr_f0=20.0)]
    orcs, ks = _pair(tubes, KAOS_FP64)
    for n, master in _steps(orcs, ks, 10, 1):
        _bitexact(orcs, ks, n, master)
    ks.close()


def test_icr_mixed_close_to_oracle():
    # bounds: alive ions equal; counts equal but for a few ions; Vperp within
    # 1e-6 of the largest (binary32 components); moments within 1e-4 (L2)
    tube = _heated(grids.config1(n_total=20000))
    orcs, ks = _pair([tube], KAOS_MIXED)
    for n, master in _steps(orcs, ks, 12, 1):
        ma, mb = orcs[0].master(), ks.master(0)
        assert np.abs(ma["nq_counts"] - mb["nq_counts"]).sum() <= 2
        assert np.abs(ma["nph_counts"] - mb["nph_counts"]).sum() <= 8
        for k in ("M0", "M1perp", "M1par", "M2", "M2perp", "M2par"):
            na = np.linalg.norm(ma[k])
            assert np.linalg.norm(ma[k] - mb[k]) <= 1e-4 * na, (n, k)
    a, b = orcs[0].particles(), ks.particles(0)
    assert np.abs(a["vperp"] - b["vperp"]).max() <= 1e-6 * np.abs(a["vperp"]).max()
    # the mixed cell word carries the Vperp cell of the next binning (the cell of
    # the kicked speed), the oracle the one of the last push: compare with the
    # cell of the oracle's current Vperp
    GL, GH = tube.arrays["vperpGL"], tube.arrays["vperpGH"]
    nxt = np.searchsorted(GH, a["vperp"], side="left") + 1
    on = (a["vperp"] >= GL[0]) & (a["vperp"] <= GH[-1])
    assert np.mean(b["vperpind"][on] == nxt[on]) > 0.999
    ks.close()
