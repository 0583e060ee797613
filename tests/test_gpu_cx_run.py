"""The full-physics configuration at a small size: H+ and O+ as two tubes of
one context on the 3-D grid, ICR heating and ion-neutral charge exchange
together, ions seeded on the device, FP64, through the C-ABI.

- Stepped eagerly (kaos_step): bit-exact against one oracle per tube (the
  heating and collision streams keyed by the tube index) after every step:
  ions, ENAs, and at master steps the 3-D counts and moments, the collision
  counts, the ENA counts, escapes and moments.
- Replayed as captured graphs (kaos_run, autotuner off so that every step
  after the first is a graph replay): the same state after the last step
  as the eager run, bit for bit.
The cross-section scale gives ~100 conversions over the 12 steps.
"""
import numpy as np
import pytest

import oracle
from kaos_amd import grids
from kaos_amd.solver import KineticSolver, KAOS_FP64, KAOS_FLAG_NO_AUTOTUNE

pytestmark = pytest.mark.gpu

PART = ("x", "y", "z", "Vx", "Vy", "Vz", "vperp", "aea", "ag", "qind", "vparind", "vperpind")
ENA = ("x", "y", "z", "Vx", "Vy", "Vz", "id", "qind")
M3 = ("n3_counts", "M0", "M1perp1", "M1perp2", "M1par", "M2", "M2perp1", "M2perp2", "M2par")
ENAM = ("counts", "M0", "M1P", "M1Q", "M1PHI", "M2", "M2P", "M2Q", "M2PHI")
NSTEPS, NDATFAC, SEED = 12, 2, 1


def _tubes():
    return grids.config5_tubes(n_h=2000, n_o=6000, nv1=8, nv2=8, sigma_scale=1e-8, ena_grid=(4, 8, 4))


def _caps(tubes):
    return [int(t.cell_counts.sum()) + (NSTEPS + 2) * (t.n_inj_lb + t.n_inj_ub) + 512 for t in tubes]


def _solver(tubes, flags=0):
    ks = KineticSolver(tubes, precision=KAOS_FP64, ndatfac=NDATFAC, seed=SEED, capacity=_caps(tubes), flags=flags)
    for i, t in enumerate(tubes):
        ks.seed(i, t.cell_counts, np.full(t.nq, t.Ts_lb), 700 + i)
    return ks


def _same(a, b, keys, what):
    for k in keys:
        assert len(a[k]) == len(b[k]), (what, k, len(a[k]), len(b[k]))
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (what, k)


def _state(ks, i):
    return ks.particles(i), ks.enas(i), ks.master3(i), ks.cx(i), ks.ena_master(i), ks.ena_moments(i)


def _compare(sa, sb, what):
    pa, ea, ma, ca, xa, za = sa
    pb, eb, mb, cb, xb, zb = sb
    _same(pa, pb, PART, what + " ions")
    _same(ea, eb, ENA, what + " ENAs")
    _same(ma, mb, M3, what + " 3-D moments")
    _same(ca, cb, ("sigma", "nu", "poi"), what + " collision counts")
    assert np.array_equal(xa["nq_ena"], xb["nq_ena"]) and np.array_equal(xa["esc"], xb["esc"]), what
    _same(za, zb, ENAM, what + " ENA moments")


def test_two_species_heating_and_charge_exchange_eager_bitexact():
    tubes = _tubes()
    print(f"two species, ICR + charge exchange, fp64 eager: {[int(t.cell_counts.sum()) for t in tubes]} ions",
          flush=True)
    orcs = []
    for i, t in enumerate(tubes):
        s = oracle.seed_ions(t, t.cell_counts, np.full(t.nq, t.Ts_lb), 700 + i, tube_index=i)
        o = oracle.Oracle(t, s["q0"], s["vpar"], s["vperp"], ndatfac=NDATFAC, seed=SEED,
                          capacity=_caps(tubes)[i], tube_index=i)
        o.set_vperp(s["v1"], s["v2"])
        orcs.append(o)
    ks = _solver(tubes)
    for n in range(1, NSTEPS + 1):
        for o in orcs:
            o.step(n)
        out = ks.step(n)
        assert out.status == 0 and out.n_alive == sum(o.count() for o in orcs), n
        for i, o in enumerate(orcs):
            _same(o.particles(), ks.particles(i), PART, f"tube {i} ions step {n}")
            _same(o.enas(), ks.enas(i), ENA, f"tube {i} ENAs step {n}")
            if n == 1 or n % NDATFAC == 0:
                _same(o.master3(), ks.master3(i), M3, f"tube {i} 3-D master {n}")
                _same(o.cx(), ks.cx(i), ("sigma", "nu", "poi"), f"tube {i} collision counts {n}")
                ea, eb = o.ena_master(), ks.ena_master(i)
                assert np.array_equal(ea["nq_ena"], eb["nq_ena"]) and np.array_equal(ea["esc"], eb["esc"]), n
                _same(o.ena_moments(), ks.ena_moments(i), ENAM, f"tube {i} ENA moments {n}")
    assert sum(len(ks.enas(i)["id"]) for i in range(len(tubes))) > 50
    ks.close()
    for o in orcs:
        o.close()


def test_two_species_graph_replay_equals_eager():
    tubes = _tubes()
    print("two species, ICR + charge exchange, fp64: kaos_run graph replay against eager steps", flush=True)
    ka = _solver(tubes)
    for n in range(1, NSTEPS + 1):
        assert ka.step(n).status == 0, n
    kb = _solver(tubes, flags=KAOS_FLAG_NO_AUTOTUNE)
    assert kb.step(1).status == 0
    assert kb.run(2, NSTEPS - 1) > 0
    for i in range(len(tubes)):
        _compare(_state(ka, i), _state(kb, i), f"tube {i}")
    assert sum(len(kb.enas(i)["id"]) for i in range(len(tubes))) > 50
    ka.close()
    kb.close()
