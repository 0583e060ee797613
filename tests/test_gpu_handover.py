"""Mid-run handover of the ions and restart files, through the C-ABI.

* kaos_download_particles (int64 Qindk1/Vparindk1/Vperpindk1) +
  kaos_download_vperp -> kaos_upload_state into a fresh context,
  kaos_upload_moments (the last master step's moments, which the next
  master step's fields filter) and kaos_resume(k).  The Cartesian triple of a
  field-line point does not fix its radius to the last bit (the library reads
  it as KineticUpdateASub does, r = |(x, y, z)|), so the bar after the handover
  is the per-step bar against the reference push: every count and cell index
  equal, and every per-ion value, moment and field within 1e-12 of the
  largest value (positions differ in the last bits; the interpolated fields
  AEAmag/AGmag carry that through the steep cell-to-cell field).
* kaos_save_state -> kaos_load_state into a fresh context: bit for bit the
  uninterrupted run (every per-ion value and index, every master output), in
  both precision forms, on the 2-D and the 3-D grid, with ICR heating.
* the Fortran driver does the same through iso_c_binding
  (KAOS_HANDOVER_STEP, KAOS_CHECKPOINT).
"""
import os
import subprocess

import numpy as np
import pytest

from kaos_amd import grids
from kaos_amd import fortran as kf
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qk4", "qind", "vparind", "vperpind")
MASTER = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA", "EG", "esc")


def _solver(tube, prec, ndatfac, cap, seed=7):
    return KineticSolver([tube], precision=prec, ndatfac=ndatfac, seed=seed, capacity=[cap])


def _tube(n=20000, south=False, icr=False):
    t = grids.dipole_tube(nq=48, nvpar=40, nvperp=30, n_total=n, ratio=1.06, south=south)
    t.n_inj_ub = max(t.n_inj_ub, 40)
    if icr:
        t.icr = 1
    return t


def _start(tube, prec, ndatfac, cap):
    q0, v0, w0 = grids.initial_ions(tube, seed=5)
    ks = _solver(tube, prec, ndatfac, cap)
    ks.upload(0, q0, v0, w0)
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * np.pi, len(w0))
    ks.upload_vperp(0, w0 * np.cos(ph), w0 * np.sin(ph))
    return ks


@pytest.mark.parametrize("k", [6, 7])
def test_fp64_handover_through_reference_arrays(k):
    """Hand over after a master step (k = 6, ndatfac 3) and after a non-master
    step (k = 7); run to step 15 in both and compare every step."""
    print(f"handover fp64 after step {k}", flush=True)
    t = _tube()
    cap = 60000
    A = _start(t, KAOS_FP64, 3, cap)
    B0 = _start(t, KAOS_FP64, 3, cap)
    for n in range(1, k + 1):
        A.step(n)
        B0.step(n)
    P = B0.particles(0, vperp12=True)
    M = B0.master(0)
    assert P["qind"].dtype == np.int64 and np.any(P["qind"] > 1)
    B0.close()
    B = _solver(t, KAOS_FP64, 3, cap)
    B.upload_state(0, P)
    B.upload_moments(0, M)
    B.resume(k)
    # the uploaded state reads back as it was handed over, to rounding
    Q = B.particles(0, vperp12=True)
    for key in ("qind", "vparind", "vperpind", "vperp1", "vperp2", "aea", "ag", "qk4"):
        assert np.array_equal(P[key], Q[key]), key
    for key in ("x", "y", "z", "Vx", "Vy", "Vz"):
        assert np.max(np.abs(P[key] - Q[key])) <= 1e-15 * np.abs(P[key]).max(), key
    for n in range(k + 1, 16):
        A.step(n)
        B.step(n)
        a, b = A.particles(0), B.particles(0)
        assert len(a["x"]) == len(b["x"]), n
        for key in ("qind", "vparind", "vperpind"):
            assert np.array_equal(a[key], b[key]), (n, key)
        for key in ("x", "y", "z", "Vx", "Vy", "Vz", "aea", "ag", "qk4"):
            scale = max(np.abs(a[key]).max(), 1e-300)
            assert np.max(np.abs(a[key] - b[key])) <= 1e-12 * scale, (n, key)
        if n % 3 == 0:
            ma, mb = A.master(0), B.master(0)
            for key in ("nq_counts", "nph_counts", "esc"):
                assert np.array_equal(ma[key], mb[key]), (n, key)
            for key in ("M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA"):
                scale = max(np.abs(ma[key]).max(), 1e-300)
                assert np.max(np.abs(ma[key] - mb[key])) <= 1e-12 * scale, (n, key)
    A.close()
    B.close()


def _bitexact(A, B, n, master, v3d=False):
    a, b = A.particles(0, vperp12=True), B.particles(0, vperp12=True)
    for key in PART + ("vperp1", "vperp2"):
        assert np.array_equal(a[key], b[key]), (n, key)
    if master:
        ma, mb = A.master(0), B.master(0)
        for key in MASTER:
            assert np.array_equal(ma[key], mb[key]), (n, key)
        if v3d:
            m3a, m3b = A.master3(0), B.master3(0)
            for key in m3a:
                assert np.array_equal(m3a[key], m3b[key]), (n, key)


@pytest.mark.parametrize("prec,v3d,icr", [(KAOS_FP64, False, False), (KAOS_MIXED, False, False),
                                          (KAOS_FP64, True, False), (KAOS_MIXED, True, False),
                                          (KAOS_FP64, False, True), (KAOS_MIXED, False, True)])
def test_checkpoint_restart_is_bitexact(tmp_path, prec, v3d, icr):
    print(f"checkpoint: precision {prec} 3-D {v3d} ICR {icr}", flush=True)
    t = _tube(south=v3d, icr=icr)
    if v3d:
        grids.with_vgrid3(t, 12, 10, vperp_max=6.0e3)
    cap = 60000
    A = _start(t, prec, 2, cap)
    S = _start(t, prec, 2, cap)
    k = 9
    for n in range(1, k + 1):
        A.step(n)
        S.step(n)
    path = tmp_path / "kaos.ckpt"
    S.save_state(path)
    S.close()
    B = _solver(t, prec, 2, cap)
    B.load_state(path)
    _bitexact(A, B, k, False, v3d)
    for n in range(k + 1, k + 8):
        A.step(n)
        B.step(n)
        _bitexact(A, B, n, n % 2 == 0, v3d)
    A.close()
    B.close()


def test_checkpoint_and_handover_errors(tmp_path):
    t = _tube(n=3000)
    A = _start(t, KAOS_FP64, 1, 10000)
    with pytest.raises(KaosError):               # nothing to save before the first step
        A.save_state(tmp_path / "x")
    A.step(1)
    A.step(2)
    A.save_state(tmp_path / "a.ckpt")
    # another grid: refused
    t2 = _tube(n=3000)
    t2.arrays = dict(t2.arrays)
    t2.arrays["Te"] = t2.arrays["Te"] * 1.5
    B = _solver(t2, KAOS_FP64, 1, 10000)
    with pytest.raises(KaosError) as e:
        B.load_state(tmp_path / "a.ckpt")
    assert e.value.code == 1
    # another precision form: refused
    C2 = _solver(t, KAOS_MIXED, 1, 10000)
    with pytest.raises(KaosError):
        C2.load_state(tmp_path / "a.ckpt")
    # truncated file
    raw = open(tmp_path / "a.ckpt", "rb").read()
    open(tmp_path / "t.ckpt", "wb").write(raw[: len(raw) // 2])
    D = _solver(t, KAOS_FP64, 1, 10000)
    with pytest.raises(KaosError) as e:
        D.load_state(tmp_path / "t.ckpt")
    assert e.value.code == 6
    # resume after a step, upload_state with a bad index
    P = A.particles(0)
    with pytest.raises(KaosError):
        A.resume(5)
    E = _solver(t, KAOS_FP64, 1, 10000)
    bad = dict(P)
    bad["qind"] = P["qind"].copy()
    bad["qind"][3] = t.nq + 1
    with pytest.raises(KaosError) as e:
        E.upload_state(0, bad)
    assert e.value.code == 1
    with pytest.raises(KaosError):
        E.resume(2)                                 # no ions yet
    for s in (A, B, C2, D, E):
        s.close()


@pytest.mark.parametrize("mode", ["handover", "checkpoint"])
def test_fortran_driver_handover(tmp_path, mode):
    if not os.path.exists(kf.DRIVER):
        pytest.skip("kaos_fdriver not built")
    print(f"Fortran driver {mode}", flush=True)
    t = grids.config1(n_total=20000)
    q0, v0, w0 = grids.initial_ions(t, seed=4)
    fin = tmp_path / "in.bin"
    kf.write_input(fin, t, q0, v0, w0, nsteps=14, ndatfac=3, seed=9)
    env = dict(os.environ)
    r = subprocess.run([kf.DRIVER, str(fin), str(tmp_path / "ref.bin")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    env["KAOS_HANDOVER_STEP"] = "7"
    if mode == "checkpoint":
        env["KAOS_CHECKPOINT"] = str(tmp_path / "f.ckpt")
    r = subprocess.run([kf.DRIVER, str(fin), str(tmp_path / "ho.bin")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    a = kf.read_output(tmp_path / "ref.bin", t.nq)
    b = kf.read_output(tmp_path / "ho.bin", t.nq)
    assert [x["n"] for x in a] == [x["n"] for x in b] and len(a) == 5
    for x, y in zip(a, b):
        assert x["n_alive"] == y["n_alive"]
        assert np.array_equal(x["nq_counts"], y["nq_counts"]) and np.array_equal(x["esc"], y["esc"])
        if mode == "checkpoint":
            assert np.array_equal(x["M"], y["M"]) and np.array_equal(x["EA"], y["EA"])
        else:
            for key in ("M", "EA"):
                scale = np.abs(x[key]).max()
                assert np.max(np.abs(x[key] - y[key])) <= 1e-12 * scale, (x["n"], key)
