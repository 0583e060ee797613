"""kaos_step_noalloc against kaos_step, through the C-ABI.

The no-allocation step stages the caller's injection uniforms in the
library's fixed per-GPU device array.  Bar: bit-exact (FP64) against kaos_step
on the same inputs, for two calls in a row of which the second stages fewer
uniforms than the first (so stale values of the first call sit past the
second call's data), repeated over several steps, then the seeded stream and
the size limit.
"""
import numpy as np
import pytest

from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KaosError, KAOS_FP64, KAOS_MIXED

pytestmark = pytest.mark.gpu

PART = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qk4", "qind", "vparind", "vperpind")
MASTER = ("nq_counts", "nph_counts", "NqRT", "M0", "M1perp", "M1par", "M2", "M2perp", "M2par", "EA", "esc")


def _tube(n_lb, n_ub, n=15000):
    t = grids.dipole_tube(nq=40, nvpar=32, nvperp=24, n_total=n, ratio=1.07)
    t.n_inj_lb, t.n_inj_ub = n_lb, n_ub
    return t


def _pair(t, prec=KAOS_FP64):
    q0, v0, w0 = grids.initial_ions(t, seed=21)
    cap = len(q0) + 12 * (t.n_inj_lb + t.n_inj_ub) + 1024
    out = []
    for _ in range(2):
        ks = KineticSolver([t], precision=prec, ndatfac=1, seed=13, capacity=[cap])
        ks.upload(0, q0, v0, w0)
        out.append(ks)
    return out


def _same(a, b, n):
    pa, pb = a.particles(0), b.particles(0)
    for k in PART:
        assert np.array_equal(pa[k], pb[k]), (n, k)
    ma, mb = a.master(0), b.master(0)
    for k in MASTER:
        assert np.array_equal(ma[k], mb[k]), (n, k)


@pytest.mark.parametrize("prec", [KAOS_FP64, KAOS_MIXED])
def test_noalloc_step_equals_kaos_step(prec):
    print(f"no-allocation step, precision {prec}", flush=True)
    big = _tube(700, 200)          # 5400 uniforms per master step
    small = _tube(90, 30)          # 720 uniforms: the second call stages less
    A_big, B_big = _pair(big, prec)
    A_small, B_small = _pair(small, prec)
    rng = np.random.default_rng(8)
    for n in range(1, 8):
        u_big = 1.0 - rng.random(6 * 900)
        u_small = 1.0 - rng.random(6 * 120)
        ob = A_big.step(n, u_big)
        nb = B_big.step(n, u_big, noalloc=True)          # first call: larger input
        os_ = A_small.step(n, u_small)
        ns = B_small.step(n, u_small, noalloc=True)      # second call: smaller input
        for o, p in ((ob, nb), (os_, ns)):
            assert (o.n_alive, o.n_injected, o.master, o.status) == (p.n_alive, p.n_injected, p.master, p.status)
        _same(A_big, B_big, n)
        _same(A_small, B_small, n)
    # the seeded stream (no caller uniforms) and a mix with kaos_step on one context
    for n in range(8, 12):
        A_small.step(n)
        if n % 2:
            B_small.step(n, noalloc=True)
        else:
            B_small.step(n)
        _same(A_small, B_small, n)
    for s in (A_big, B_big, A_small, B_small):
        s.close()


def test_noalloc_step_refuses_more_than_its_scratch():
    t = _tube(60000, 6000, n=4000)         # 66000 ions per step > 65536
    ks = KineticSolver([t], precision=KAOS_FP64, ndatfac=1, seed=3, capacity=[4000 + 3 * 66000 + 1024])
    q0, v0, w0 = grids.initial_ions(t, seed=2)
    ks.upload(0, q0, v0, w0)
    ks.step(1, noalloc=True)               # step 1 injects nothing
    u = 1.0 - np.random.default_rng(1).random(6 * 66000)
    with pytest.raises(KaosError) as e:
        ks.step(2, u, noalloc=True)
    assert e.value.code == 1
    ks.step(2, u)                          # kaos_step takes it
    assert ks.count(0) > 60000
    ks.close()
