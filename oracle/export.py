"""TEST INFRASTRUCTURE ONLY.  The oracle's restatement of DataExport2Sub
(_8_0_2_DataExport2.f90:25-704): the stream-unformatted files of one flux
tube at master step nn, written from the oracle's master outputs with numpy.

Flag set (this path): FLUIDIONEXPORTflag, PHASEIONDISTRIBflag, the moment
flags and MOMENTFILTERflag on; QEXCHANGEflag, SPINUPflag, EAINERTIALflag and
EPARflag off (_ModelInput.f90); ION2VPERPflag either (the 3-D branch when
the tube has a 3-D velocity grid and master3 is given).

File names: DataExport2Sub builds "<nn>_<s>_<f>_" from I5 edits of nn, s
and f with adjustl/adjustr (:32-38), and "<nn>_<s>_<f>_<Qind>_" for the
per-cell files (:129-137); the reference's reader KAOS_Data_Plots.py:60-70
opens str(nn) + '_' + str(s) + '_' + str(f) + '_<Name>fort.bin'.  Contents:
form='unformatted', access='stream' writes the array in Fortran order with
no record markers (little-endian binary64; ndatfacGT is integer(dp), int64).
"""
import os

import numpy as np


def file_name(nn, s, f, name, Q=None):
    """<nn>_<s>_<f>_[<Q>_]<name>fort.bin (1-based integers, no padding)."""
    mid = f"{nn}_{s}_{f}_" + ("" if Q is None else f"{Q}_")
    return mid + name + "fort.bin"


def _put(path, arr, dtype="<f8"):
    np.ascontiguousarray(np.asarray(arr, dtype=dtype)).tofile(path)


def nph_rt(counts, nsnormfac, d3x):
    """NphRTp(:, :, nn) of every cell (_8_3_1_IonParticleCounts.f90:713-729):
    NphReNormRT (nsnormfac / d3xCGT), 0 where the count is 0; counts are
    [Q][Vperp][Vpar]; returns [Q] arrays in Fortran order (Vperp fastest)."""
    out = []
    for Q in range(counts.shape[0]):
        w = nsnormfac / d3x[Q]
        c = counts[Q].astype(np.float64)
        v = np.where(counts[Q] == 0, 0.0, c * w)
        out.append(np.asfortranarray(v).ravel(order="F"))
    return out


def n2perp_rt(counts3, nsnormfac, d3x):
    """N2PerpphRTp(:, :, :, nn) of every cell (_8_3_1_IonParticleCounts.f90:
    440-457): N2PerpphReNormRT (nsnormfac / d3xCGT), 0 where the count is 0;
    counts are [Q][V1][V2][Vpar]; returns [Q] arrays in Fortran order
    (Vperp1 fastest, then Vperp2, then Vpar)."""
    out = []
    for Q in range(counts3.shape[0]):
        w = nsnormfac / d3x[Q]
        v = np.where(counts3[Q] == 0, 0.0, counts3[Q].astype(np.float64) * w)
        out.append(np.asfortranarray(v).ravel(order="F"))
    return out


def export_master(directory, tube, master, n, ndatfac, theta, Ts, ns0, phi, s=1, f=1, master3=None):
    """Write the files of master step n (n = 1, or a master step) of one tube
    from the oracle's master() outputs (and master3() on the 3-D grid);
    returns nn."""
    v3d = master3 is not None
    nn = 1 if n == 1 else n // ndatfac + 1
    nq = tube.nq
    a = tube.arrays
    put = lambda name, v, dt="<f8": _put(os.path.join(directory, file_name(nn, s, f, name)), v, dt)
    # :29-107
    put("TimeT", [0.0 if n == 1 else float(n) * tube.h])      # _8_KineticSolver.f90:178-184
    put("ndatfacGT", [ndatfac], "<i8")
    put("ns0GT", [ns0])
    put("qGCGT", a["qGC"])
    put("pGCGT", np.full(nq, tube.p))
    put("rGCGT", a["rGC"])
    put("phiGCGT", np.full(nq, phi))
    put("thetaGCGT", theta)
    put("d3xCGT", a["d3xC"])
    put("TsGT", Ts)
    if v3d:
        # :124-165: TeNT(nn) (constant: dNTe = 0, _8_KineticSolver.f90:66-84),
        # N2PerpphRTp(:, :, :, nn), d3vCTp(:, :, :) of every cell
        d3 = np.asfortranarray(a["d3v3"].reshape(tube.nv1, tube.nv2, tube.nvpar)).ravel(order="F")
        for Q, v in enumerate(n2perp_rt(master3["n3_counts"], tube.nsnormfac, a["d3xC"])):
            _put(os.path.join(directory, file_name(nn, s, f, "TeNT", Q + 1)), [a["Te"][Q]])
            _put(os.path.join(directory, file_name(nn, s, f, "N2PerpphRT", Q + 1)), v)
            _put(os.path.join(directory, file_name(nn, s, f, "d3vCTp", Q + 1)), d3)
    else:
        # :167-191
        for Q, v in enumerate(nph_rt(master["nph_counts"], tube.nsnormfac, a["d3xC"])):
            _put(os.path.join(directory, file_name(nn, s, f, "NphRT", Q + 1)), v)
    # :225-510 (2-D branch) and :359-410 (filter: M1Perp1/M1Perp2/M2Par are set
    # only by the charge-exchange collision frequency, _8_6_6_1:50-92)
    z = np.zeros(nq)
    put("M0phRT", master["M0"])
    if v3d:                                                    # :279-305
        put("M1Perp1phRT", master3["M1perp1"])
        put("M1Perp2phRT", master3["M1perp2"])
    else:
        put("M1PerpphRT", master["M1perp"])
    put("M1ParphRT", master["M1par"])
    put("M0FiltAvrgRT", master["M0F"])
    put("M1Perp1FiltAvrgRT", z)
    put("M1Perp2FiltAvrgRT", z)
    put("M1ParFiltAvrgRT", master["M1F"])
    put("M2ParFiltAvrgRT", z)
    put("M2phRT", master["M2"])
    if v3d:                                                    # :468-500
        put("M2Perp1phRT", master3["M2perp1"])
        put("M2Perp2phRT", master3["M2perp2"])
    else:
        put("M2PerpphRT", master["M2perp"])
    put("M2ParphRT", master["M2par"])
    # :645-704; EAINERTIALflag off (_8_7_AmbipolarEfield.f90:153-155, 1296-1299),
    # EPARflag off (_8_8_PotentialStructure.f90:46-96)
    put("EAInertialRT", z)
    put("EAPressureRT", master["EA"])
    put("EAmagRT", 0.0 + master["EA"])
    put("EGmagRT", master["EG"])
    put("EPmagRT", z)
    put("PhiParRT", z)
    return nn


KB = 1.38e-23                       # kB = 138d-25 (_KineticMainParams.f90)


def spinup_name(s, f, name):
    """<s>_<f>_<name>fort.bin (_8_0_2_DataExport2.f90:706-712)."""
    return f"{s}_{f}_{name}fort.bin"


def export_spinup(directory, tube, master, master3, nq_lb, nq_ub, nsnorm_clb, nsnorm_cub,
                  lb_replenish, ub_replenish, s=1, f=1):
    """The spin-up files (SPINUPflag = 1, _8_0_2_DataExport2.f90:704-802) of
    the last master step: DensityOutputRT = M0phRT and TemperatureOutputRT =
    (2/3) M2ParphRT/kB + (2/3) M2Perp1phRT/kB + (2/3) M2Perp2phRT/kB
    (_8_6_6_SecondParIonMoment.f90:659-673, left to right); the EA outputs of
    _8_7_AmbipolarEfield.f90:1419-1421 (EAINERTIALflag off: 0); the caller's
    NqLBT, NqUBT, nsnormC(L/U)BGT and replenish flags."""
    put = lambda name, v, dt="<f8": _put(os.path.join(directory, spinup_name(s, f, name)), v, dt)
    c23 = 2.0 / 3.0
    T = np.array([((c23 * a) / KB + (c23 * b) / KB) + (c23 * c) / KB
                  for a, b, c in zip(master3["M2par"], master3["M2perp1"], master3["M2perp2"])])
    z = np.zeros(tube.nq)
    put("NqLBT", [nq_lb], "<i8")
    put("NqUBT", [nq_ub], "<i8")
    put("DensityOutputRT", master3["M0"])
    put("TemperatureOutputRT", T)
    put("EAInertialOutputRT", z)
    put("EAPressureOutputRT", master["EA"])
    put("EAmagOutputRT", 0.0 + master["EA"])
    put("nsnormCLBOutput", [nsnorm_clb])
    put("nsnormCUBOutput", [nsnorm_cub])
    put("LBREPLENISHflagTOutput", [lb_replenish], "<i8")
    put("UBREPLENISHflagTOutput", [ub_replenish], "<i8")
    return T
