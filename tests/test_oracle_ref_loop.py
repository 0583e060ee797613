"""Pins of the oracle against the reference's own set-up and step loop (CPU).

oracle/_ref/kaos_ref_loop is the reference's program (every subroutine of its
set-up and of KineticSolverSub compiled unmodified from its sources, 2 MPI
ranks) with the harness's input module (one O+ tube on L = 4, 64 config cells,
12 x 12 x 12 (Vperp1, Vperp2, Vpar) cells, ndatfac 3) and a logged, seeded
uniform stream in place of random_number.  The oracle is run on the inputs
the reference generated and compared with what the reference computed:

* the reference's grid is one the library accepts (contiguous edges);
* every initial ion's (Vpar, Vperp1, Vperp2) is the oracle's Maxwell-Boltzmann
  projection of the reference's own six logged uniforms at the ion's point and
  cell temperature (_6_VelocityDistribution.f90, _6_1_MBVelocityDistribution.f90);
* step 1's 3-D phase-space counts: the reference's k0 cells come from its
  q-edge arrays read at master row nn = NNtT + 2, one row past their end
  (KineticSolverASub, _8_0_0_1_KineticSolverA.f90:52-171, runs after
  BoundaryConditionsSub's loop over nn has left nn there), which in the
  column-major arrays is row 1 of the next cell: every ion is counted one
  config cell below its own (cell 1 keeps its ions and takes cell 2's, the
  top cell is empty) while its three velocity cells are right.  The oracle's
  counts reproduce that exactly once shifted the same way;
* with the ions placed so that the oracle's step-1 counts are the reference's,
  its eight 3-D moments of step 1 and the gravity field equal the reference's
  exported M0phRT .. M2ParphRT and EGmagRT within 1e-12.

Master steps 2 to 5 are compared on oracle/_ref/kaos_ref_loop_eamb0 (the same
program with EAMBflag = 0): with the ambipolar field on, the reference's field
of master step 1 reads row nn - 1 = 0 of its moment arrays (out of range,
_8_7_AmbipolarEfield.f90:40-82), so its first pushes see an undefined field.
Three properties of the compiled reference's sequence are measured there;
the first two are fed to the oracle and the third is the oracle's own rule:

* step 1 clamps q0 against the edge arrays read at that same out-of-range row
  (KSA:55-66; in the column-major arrays, the next cell's edges): every ion
  of cell 1 takes q0 = qGL of cell 2, and GravfieldSub interpolates the first
  AGmagN there (_8_7_2_Gravfield.f90:132-300) -- Oracle.set_q0;
* the ions advance at step 1 and at master steps only: the state the
  reference's loop leaves after 6 steps is the reference's own push applied
  at n = 1, 3 and 6, bit for bit, and not at every step: the non-master
  branch of KineticSolverBSub matches only for nn >= 2 (KSB:804-808), and the
  `exit nnloopKSnonstat` that ends the loop body (KSB:1033) leaves the loop
  during its first pass, nn = 1 -- so the oracle is stepped at n = 1, 3, 6,
  9, 12;
* an injected ion's first push runs without gravity: its slot's AGmagN and
  qk4 are never written before that push (DataTypeReAllocA copies only the
  old ions, _8_0_6_DataTypeReAllocA.f90:259-313, and no cell of GravfieldSub's
  scan holds an unset qk4, _8_7_2_Gravfield.f90:305-420), and the compiled
  reference reads 0 there for the batches these master steps export (the
  injection cell's field instead bins a few ions per 1e4 differently from
  master step 4 on).

With these, the 3-D phase-space counts of master steps 2 to 5 are equal, the
eight moments agree within 1e-13 and the gravity field and the lower-boundary
escape flux (ions removed at the next master step, BC:158-256) equal the
reference's: the cadence, removal lag, injection at the boundary cell with the
logged uniforms and the field lag of the oracle are the reference's.
"""
import math
import os

import numpy as np
import pytest

import oracle
import oracle.ref_loop as rl
from oracle import ptr

NT = 9                    # three master steps (ndatfac 3)
NSNORMFAC = 3.85e17       # about 2e4 ions


@pytest.fixture(scope="module")
def ref(tmp_path_factory):
    if not rl.available():
        if rl.build() is None or not rl.available():
            pytest.skip("reference sources or MPI absent: the reference loop cannot be built or run")
    wd = str(tmp_path_factory.mktemp("kaos_ref_loop"))
    timing = rl.run(wd, NT, NSNORMFAC, ranks=2, seed=3)
    inp = rl.load_inputs(wd)
    return dict(wd=wd, timing=timing, inp=inp)


def test_reference_run_and_grid(ref):
    tm, inp = ref["timing"], ref["inp"]
    t = inp["tube"]
    assert tm["ranks"] == 2 and tm["steps"] == NT and tm["ndatfac"] == 3 and tm["seconds"] > 0
    assert tm["ions0"] == len(inp["ions"]["q0"]) > 10000
    a = t.arrays
    for lo, hi in (("qGL", "qGH"), ("vparGL", "vparGH"), ("v1GL", "v1GH"), ("v2GL", "v2GH")):
        assert np.array_equal(a[lo][1:], a[hi][:-1]), lo
    assert (t.nq, t.nv1, t.nv2, t.nvpar) == (64, 12, 12, 12)
    # the logged stream holds 6 uniforms per initial ion and per injected ion
    for r, u in enumerate(inp["uniforms"]):
        assert len(u) == 6 * (inp["counts"][r] + (NT // 3) * inp["n_inj_rank"])


def test_initial_velocities_are_the_mb_projection_of_the_logged_uniforms(ref):
    inp = ref["inp"]
    t, I = inp["tube"], inp["ions"]
    lib = oracle.load()
    qGC = t.arrays["qGC"]
    o4 = np.zeros(4)
    worst = 0.0
    j = 0
    for r, n in enumerate(inp["counts"]):
        u = inp["uniforms"][r]
        for k in range(n):
            q = I["q0"][j]
            Q = int(np.argmin(np.abs(qGC - q)))
            assert abs(qGC[Q] - q) <= 1e-12 * abs(q)        # DensityProfileBSub: cell centres
            sigma = math.sqrt(1.38e-23 * inp["Ts"][Q] / t.ms)
            rs = lib.orc_root_r(q, t.p)
            lib.orc_mb(ptr(np.ascontiguousarray(u[6 * k:6 * k + 6])), sigma, t.cphi, t.sphi, rs, t.p, ptr(o4))
            got = np.array([I["vpar0"][j], I["v1"][j], I["v2"][j], I["vperp"][j]])
            worst = max(worst, float(np.max(np.abs(o4 - got)) / sigma))
            j += 1
    assert worst < 1e-12, worst


def _n3_of_step1(inp, shift):
    t, I = inp["tube"], inp["ions"]
    qGC = t.arrays["qGC"]
    q0 = I["q0"].copy()
    if shift:
        # the reference's k0 cell: one below the ion's own (cell 1 stays)
        Q = np.array([int(np.argmin(np.abs(qGC - q))) for q in q0])
        q0 = qGC[np.maximum(Q - 1, 0)]
    o = oracle.Oracle(t, q0, I["vpar0"], I["vperp"], ndatfac=3, seed=1, capacity=len(q0) + 4096)
    o.set_vperp(I["v1"], I["v2"])
    o.step(1)
    out = o.master3(), o.master()
    o.close()
    return out


def test_step1_counts_are_the_oracles_shifted_one_cell(ref):
    inp = ref["inp"]
    R = rl.read_master(ref["wd"], 1, inp["tube"])
    assert np.max(np.abs(R["n3_real"] - R["n3_counts"])) < 1e-9     # integer counts x nsnormfac / d3x
    ours = _n3_of_step1(inp, shift=False)[0]["n3_counts"]
    ref3 = R["n3_counts"]
    assert np.array_equal(ref3[0], ours[0] + ours[1])
    assert np.array_equal(ref3[1:-1], ours[2:])
    assert not ref3[-1].any()
    # the three velocity cells of every ion agree
    for ax in ((0, 2, 3), (0, 1, 3), (0, 1, 2)):
        assert np.array_equal(ref3.sum(axis=ax), ours.sum(axis=ax))


def test_step1_moments_and_gravity_equal_the_references(ref):
    inp = ref["inp"]
    t = inp["tube"]
    R = rl.read_master(ref["wd"], 1, t)
    m3, m = _n3_of_step1(inp, shift=True)
    assert np.array_equal(m3["n3_counts"], R["n3_counts"])
    for key in oracle.MOMENTS3:
        scale = max(np.abs(R[key]).max(), 1e-300)
        assert np.max(np.abs(m3[key] - R[key])) <= 1e-12 * scale, key
    assert np.max(np.abs(m["EG"] - R["EG"])) <= 1e-12 * np.abs(R["EG"]).max()


NT_SEQ = 12               # steps of the sequencing run (master steps at 3, 6, 9, 12)


@pytest.fixture(scope="module")
def ref_eamb0(tmp_path_factory):
    if not rl.available(rl.EXE_EAMB0):
        if rl.build() is None or not rl.available(rl.EXE_EAMB0):
            pytest.skip("reference sources or MPI absent: the reference loop cannot be built or run")
    wd = str(tmp_path_factory.mktemp("kaos_ref_loop_eamb0"))
    rl.run(wd, NT_SEQ, NSNORMFAC, ranks=2, seed=3, exe=rl.EXE_EAMB0)
    return dict(wd=wd, inp=rl.load_inputs(wd))


def test_master_steps_2_to_5_equal_the_references(ref_eamb0):
    inp = ref_eamb0["inp"]
    t, I = inp["tube"], inp["ions"]
    t.eamb = 0
    qGL = t.arrays["qGL"]
    # step 1's q0: clamped at cell 2's lower edge (the row past the end of cell 1's)
    q0s = np.minimum(I["q0"], qGL[1]) if qGL[0] > 0 else np.maximum(I["q0"], qGL[1])
    assert np.sum(q0s != I["q0"]) > 0
    cap = len(I["q0"]) + 8 * inp["n_inj_rank"] * inp["ranks"] + 1024
    o = oracle.Oracle(t, I["q0"], I["vpar0"], I["vperp"], ndatfac=3, seed=1, capacity=cap)
    o.set_vperp(I["v1"], I["v2"])
    o.set_q0(q0s)
    o.step(1)
    for n in (3, 6, 9, 12):
        o.step(n, rl.injection_uniforms(inp, n // 3))
        nn = n // 3 + 1
        R = rl.read_master(ref_eamb0["wd"], nn, t)
        m3, m = o.master3(), o.master()
        assert np.array_equal(m3["n3_counts"], R["n3_counts"]), nn
        for key in oracle.MOMENTS3:
            scale = max(np.abs(R[key]).max(), 1e-300)
            assert np.max(np.abs(m3[key] - R[key])) <= 1e-13 * scale, (nn, key)
        assert np.max(np.abs(m["EG"] - R["EG"])) <= 1e-12 * np.abs(R["EG"]).max(), nn
        assert abs(m["esc"][2] - R["lb_out"][0]) <= 1e-12 * max(abs(R["lb_out"][0]), 1.0), nn
    assert m["esc"][0] > 0                       # ions left through the lower boundary by master step 5


def test_ambipolar_field_of_master_steps_3_and_4_equals_the_references(ref):
    """The reference's loop with its ambipolar field on.  The fields of master
    steps 1 and 2 come from moments the oracle does not reproduce (row nn - 1 =
    0 of the moment arrays, out of range, at step 1; step 1's counts one cell
    low at step 2), so both are fed from the reference's exports
    (Oracle.set_field_input).  From there the oracle's own field -- the
    stencils of M0, the LB override and the M1par filter (_8_7_AmbipolarEfield.
    f90) -- equals the reference's at master steps 3 and 4, and the counts,
    moments and LB escape flux of master steps 2 and 3 are the reference's."""
    inp = ref["inp"]
    t, I = inp["tube"], inp["ions"]
    assert t.eamb == 1
    qGL = t.arrays["qGL"]
    q0s = np.minimum(I["q0"], qGL[1]) if qGL[0] > 0 else np.maximum(I["q0"], qGL[1])
    cap = len(I["q0"]) + 4 * inp["n_inj_rank"] * inp["ranks"] + 1024
    R = {nn: rl.read_master(ref["wd"], nn, t) for nn in (1, 2, 3, 4)}
    o = oracle.Oracle(t, I["q0"], I["vpar0"], I["vperp"], ndatfac=3, seed=1, capacity=cap)
    o.set_vperp(I["v1"], I["v2"])
    o.set_q0(q0s)
    o.set_field_input(np.ascontiguousarray(R[1]["EA"]), 1)
    o.step(1)
    o.set_field_input(np.ascontiguousarray(R[2]["EA"]), 2)
    for n in (3, 6, 9):
        nn = n // 3 + 1
        if n == 6:
            o.set_field_input(None, 0)                  # the oracle's own field from here on
        o.step(n, rl.injection_uniforms(inp, n // 3))
        m3, m = o.master3(), o.master()
        if nn >= 3:
            scale = np.abs(R[nn]["EA"]).max()
            assert scale > 0 and np.max(np.abs(m["EA"] - R[nn]["EA"])) <= 1e-13 * scale, nn
        if nn <= 3:
            assert np.array_equal(m3["n3_counts"], R[nn]["n3_counts"]), nn
            for key in oracle.MOMENTS3:
                sc = max(np.abs(R[nn][key]).max(), 1e-300)
                assert np.max(np.abs(m3[key] - R[nn][key])) <= 1e-13 * sc, (nn, key)
            assert abs(m["esc"][2] - R[nn]["lb_out"][0]) <= 1e-12 * max(abs(R[nn]["lb_out"][0]), 1.0), nn
    # the field of master step 3 is not the fed field of step 2 carried over:
    # the oracle computed it, so the pin is not vacuous
    assert np.max(np.abs(R[3]["EA"] - R[2]["EA"])) > 1e-3 * np.abs(R[3]["EA"]).max()
