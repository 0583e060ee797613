"""Pins of the oracle's initial ions (CPU, no GPU): per-cell macro-ion counts
(_0_1_1_ConfigGridGeneratorB.f90:83-85, _2_1_DensityProfileA1.f90:45), ions
at cell centres (_4_DensityProfileB.f90:40-46), root solve
(DipolePolynomialSolverSub) and Maxwell-Boltzmann velocities
(_6_VelocityDistribution.f90:60-85) from the seeded Philox stream.

The reference draws its uniforms from Fortran's random_number, which no
restatement can reproduce; the stream here is pinned against numpy's Philox
and the velocities by the rotation invariant and the Maxwellian's moments.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import ptr
from kaos_amd import grids, solver

from test_oracle_pins import _ref_complex_root


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def _int_exact(x):
    """Fortran INT (assignment to an integer) of a binary64 value: truncation
    toward zero, in exact rational arithmetic."""
    f = Fraction(x)
    n = math.floor(abs(f))
    return n if f >= 0 else -n


def test_cell_count_truncates_toward_zero(lib):
    """nsnormC0 is an integer array (_KineticMainParams.f90:64): the density
    times d3x / nsnormfac is truncated on assignment; the reference's initial
    ions per cell are these counts (pinned against the reference's own run in
    test_oracle_ref_loop.py)."""
    for dens, fac, d3x, want in ((5.0, 2.0, 1.0, 2), (1.0, 2.0, 1.0, 0), (3.0, 2.0, 1.0, 1),
                                 (7.0, 2.0, 1.0, 3), (0.0, 1.0, 5.0, 0), (1.0, 4.0, 1.0, 0),
                                 (0.9999999999999999, 1.0, 1.0, 0), (3.0, 1.0, 1.0, 3), (2.9999999999999996, 1.0, 1.0, 2),
                                 (-2.5, 1.0, 1.0, -2), (-2.9999999999999996, 1.0, 1.0, -2)):
        assert lib.orc_cell_count(dens, d3x, fac) == want, (dens, fac, d3x)
    rng = np.random.default_rng(3)
    dens = 10.0 ** rng.uniform(6, 12, 4000)
    d3x = 10.0 ** rng.uniform(12, 17, 4000)
    fac = 10.0 ** rng.uniform(14, 20, 4000)
    for a, b, c in zip(dens, d3x, fac):
        x = (a / c) * b
        if x < 9e18:
            assert lib.orc_cell_count(a, b, c) == _int_exact(x)
    assert lib.orc_cell_count(float("nan"), 1.0, 1.0) == -1
    assert lib.orc_cell_count(1e30, 1e10, 1.0) == -1


def test_library_cell_counts_equal_the_oracle(lib):
    """kaos_cell_counts is host arithmetic of the HIP library (no device call)."""
    t = grids.config1(n_total=50000)
    n = 1e10 * np.exp(-t.s_centre / 5.0e5)
    got = solver.cell_counts(n, t.arrays["d3xC"], t.nsnormfac)
    want = [lib.orc_cell_count(a, b, t.nsnormfac) for a, b in zip(n, t.arrays["d3xC"])]
    assert got.tolist() == want
    assert got[0] > 0 and got.sum() > 1000
    with pytest.raises(solver.KaosError):
        solver.cell_counts(n, t.arrays["d3xC"], 0.0)
    bad = n.copy()
    bad[3] = float("inf")
    with pytest.raises(solver.KaosError):
        solver.cell_counts(bad, t.arrays["d3xC"], t.nsnormfac)


def test_seed_stream_is_numpy_philox(lib):
    """Uniform k of ion i: ((w >> 11) + 1) 2^-53 of Philox4x64-10 word k, key
    (seed, "KAOS"), counter (i, 0, tube, "INI" + block)."""
    u = np.zeros(6)
    for seed, tube, i in ((1, 0, 0), (0xFEEDFACE, 3, 12345678), (2 ** 64 - 1, 511, 2 ** 40)):
        lib.orc_seed_uniforms(seed, tube, i, ptr(u))
        words = []
        for blk in (0x494E4900, 0x494E4901):
            # numpy increments the counter before each block: start one below
            ctr = np.array([(i - 1) % 2 ** 64, 0 if i > 0 else 2 ** 64 - 1, tube, blk], dtype=np.uint64)
            if i == 0:
                ctr[2] = tube - 1 if tube > 0 else 2 ** 64 - 1
                if tube == 0:
                    ctr[3] = blk - 1
            bg = np.random.Philox(counter=ctr, key=np.array([seed, 0x4B414F53], dtype=np.uint64))
            words += [int(w) for w in bg.random_raw(4)]
        want = [((w >> 11) + 1) * 2.0 ** -53 for w in words[:6]]
        assert u.tolist() == want, (seed, tube, i)
    # apart from the injection stream of the same numbers
    v = np.zeros(6)
    lib.orc_seed_uniforms(9, 0, 5, ptr(u))
    lib.orc_inject_uniforms(9, 0, 0, 5, ptr(v))
    assert not np.array_equal(u, v)


@pytest.mark.parametrize("south", [False, True])
def test_seeded_ions_at_cell_centres_with_maxwellian_velocities(lib, south):
    t = grids.dipole_tube(nq=48, nvpar=32, nvperp=32, n_total=60000, ratio=1.06, south=south)
    counts = t.cell_counts.copy()
    counts[5] = 0                                   # empty cells inside and at the top
    counts[-3:] = 0
    temp = np.linspace(1500.0, 9000.0, t.nq)
    seed = 424242
    s = oracle.seed_ions(t, counts, temp, seed, tube_index=2)
    n = int(counts.sum())
    cell = np.repeat(np.arange(t.nq), counts)
    # positions: the cell centre, radius of the reference's complex Ferrari root
    assert np.array_equal(s["q0"], t.arrays["qGC"][cell])
    for Q in np.unique(cell)[::7]:
        j = np.searchsorted(cell, Q)
        r = _ref_complex_root(t.arrays["qGC"][Q], t.p)
        assert abs(abs(s["rs"][j]) - r) <= 1e-12 * r
        assert math.copysign(1.0, s["rs"][j]) == math.copysign(1.0, t.arrays["qGC"][Q])
    # velocities: a rotation of sigma (g1, g2, g3), g Box-Muller of the stream
    u = np.zeros(6)
    sig = np.sqrt((1.38e-23 * temp) / t.ms)
    for j in range(0, n, 997):
        lib.orc_seed_uniforms(seed, 2, j, ptr(u))
        g = [lib.orc_gauss(u[2 * k], u[2 * k + 1]) for k in range(3)]
        v2 = s["vpar"][j] ** 2 + s["v1"][j] ** 2 + s["v2"][j] ** 2
        want = sig[cell[j]] ** 2 * (g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
        assert abs(v2 - want) <= 1e-12 * want
    assert np.array_equal(s["vperp"], np.abs(np.sqrt(s["v1"] ** 2 + s["v2"] ** 2)))
    # Maxwellian moments per component, pooled over cells in units of sigma
    z = {k: s[k] / sig[cell] for k in ("vpar", "v1", "v2")}
    for k, a in z.items():
        assert abs(a.mean()) < 5.0 / math.sqrt(n), k
        assert abs(a.var() - 1.0) < 6.0 * math.sqrt(2.0 / n), k
    assert abs(np.mean(z["vpar"] * z["v1"])) < 5.0 / math.sqrt(n)
    # each populated cell's spread is its own temperature's thermal speed
    big = [Q for Q in range(t.nq) if counts[Q] >= 1500]
    assert len(big) >= 4 and temp[big[-1]] > 1.5 * temp[big[0]]
    for Q in big:
        w = (s["vpar"][cell == Q] ** 2 + s["v1"][cell == Q] ** 2 + s["v2"][cell == Q] ** 2) / 3.0
        assert abs(w.mean() / sig[Q] ** 2 - 1.0) < 5.0 * math.sqrt(2.0 / (3.0 * counts[Q])), Q
