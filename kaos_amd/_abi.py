"""ctypes view of the libkaos_hip.so C-ABI (include/kaos_hip.h).

This is the binding a Python caller uses; the Fortran driver uses the
iso_c_binding module in kaos_amd/fortran/kaos_hip_iface.f90 with the same
layouts.  The library is loaded from the package directory only: a missing or
stale build raises instead of falling back to anything else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkaos_hip.so")

KAOS_OK, KAOS_E_ARG, KAOS_E_CAPACITY, KAOS_E_DEVICE, KAOS_E_STATE, KAOS_E_NAN, KAOS_E_IO = range(7)
KAOS_FP64, KAOS_MIXED = 0, 1
KAOS_FLAG_NO_GRAPH, KAOS_FLAG_NO_AUTOTUNE, KAOS_FLAG_VPERP3D = 1, 2, 4

_dp = C.POINTER(C.c_double)


class KaosCfg(C.Structure):
    _fields_ = [("ntubes", C.c_int32), ("precision", C.c_int32), ("ndatfac", C.c_int32),
                ("device", C.c_int32), ("seed", C.c_uint64), ("reorder_every", C.c_int32),
                ("flags", C.c_int32)]


GRID_ARRAYS = ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Te",
               "vparGL", "vparGH", "vparGC", "dvpar", "vperpGL", "vperpGH", "vperpGC", "dvperp",
               "d3v")


# 3-D velocity grid: kaos_grid field name -> grids.Tube array name
GRID_ARRAYS3 = ("vperp1GL", "vperp1GH", "vperp1GC", "dvperp1", "vperp2GL", "vperp2GH", "vperp2GC", "dvperp2",
                "d3v3")
TUBE_ARRAYS3 = ("v1GL", "v1GH", "v1GC", "dv1", "v2GL", "v2GH", "v2GC", "dv2", "d3v3")

# ENA velocity grid of a charge-exchange tube (kaos_grid field name = grids.Tube array name)
ENA_ARRAYS = ("vpGL", "vpGH", "vpGC", "dvp", "vqGL", "vqGH", "vqGC", "dvq", "vphiGL", "vphiGH", "vphiGC", "dvphi",
              "ena_hh", "ena_d33v")

ICR_FIELDS = ("icr_lambda_perp", "icr_eta_lh", "icr_xi_perp1", "icr_xi_perp2", "icr_s0", "icr_f0",
              "icr_chi_perp1", "icr_chi_perp2")


class KaosGrid(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nq", "nvpar", "nvperp", "grav", "eamb", "eambsign",
                                          "filter_pt", "lbcond", "ubcond", "filter_pt_m1par")]
                + [("n_inj_lb", C.c_int64), ("n_inj_ub", C.c_int64)]
                + [(n, C.c_double) for n in ("p", "cphi", "sphi", "h", "ms", "qs", "nsnormfac",
                                             "Ts_lb", "Ts_ub")]
                + [(n, _dp) for n in GRID_ARRAYS]
                + [("icr", C.c_int32), ("pad1", C.c_int32)]
                + [(n, C.c_double) for n in ICR_FIELDS]
                + [("nvperp1", C.c_int32), ("nvperp2", C.c_int32)]
                + [(n, _dp) for n in GRID_ARRAYS3]
                + [(n, C.c_int32) for n in ("cx", "cx_filter", "cx_filter_pt", "cx_pad")]
                + [(n, C.c_double) for n in ("t_neut", "m_neut", "cx_sigma_scale")]
                + [("ns_neut", _dp), ("ena_cap", C.c_int64)]
                + [(n, C.c_int32) for n in ("nvp", "nvq", "nvphi", "ena_pad")]
                + [(n, _dp) for n in ENA_ARRAYS])


class KaosExportInfo(C.Structure):
    _fields_ = [("dir", C.c_char_p), ("species", C.c_int32), ("flux_tube", C.c_int32),
                ("ns0", C.c_double), ("phi", C.c_double), ("theta", _dp), ("Ts", _dp)]


class KaosSpinupInfo(C.Structure):
    _fields_ = [("dir", C.c_char_p), ("species", C.c_int32), ("flux_tube", C.c_int32),
                ("nq_lb", C.c_int64), ("nq_ub", C.c_int64), ("nsnorm_clb", C.c_double),
                ("nsnorm_cub", C.c_double), ("lb_replenish", C.c_int64), ("ub_replenish", C.c_int64)]


class KaosSpinupState(C.Structure):
    _fields_ = [("nq_lb", C.c_int64), ("nq_ub", C.c_int64), ("nsnorm_clb", C.c_double),
                ("nsnorm_cub", C.c_double), ("lb_replenish", C.c_int64), ("ub_replenish", C.c_int64),
                ("nq", C.c_int64)]


KAOS_FIELD_COMPUTED, KAOS_FIELD_INPUT_FIRST, KAOS_FIELD_INPUT_FIXED = 0, 1, 2


class KaosStepOut(C.Structure):
    _fields_ = [("n_alive", C.c_int64), ("n_injected", C.c_int64), ("master", C.c_int32),
                ("status", C.c_int32)]


EXPORTS = ("kaos_init", "kaos_set_grid", "kaos_upload_particles", "kaos_upload_vperp", "kaos_cell_counts",
           "kaos_gridgen_dims_of", "kaos_generate_grid", "kaos_seed_particles", "kaos_step", "kaos_step_noalloc", "kaos_run",
           "kaos_download_master", "kaos_download_master3d", "kaos_export_master", "kaos_export_spinup",
           "kaos_read_spinup", "kaos_set_field",
           "kaos_count", "kaos_download_cx", "kaos_download_enas", "kaos_download_ena_master",
           "kaos_download_ena_moments",
           "kaos_download_particles", "kaos_download_vperp", "kaos_upload_state",
           "kaos_upload_moments", "kaos_resume", "kaos_save_state", "kaos_load_state", "kaos_set_launch",
           "kaos_get_launch", "kaos_set_launch_block", "kaos_get_launch_block", "kaos_profile",
           "kaos_profile_read", "kaos_last_error", "kaos_finalize")

_lib = None


# grid set-up (kaos_generate_grid): inputs, sizes and the per-cell output arrays
GRIDGEN_CFG = ("qGL", "qGH", "qGC", "hqC", "dqC", "d3xC", "rGC", "cthGC", "ellGC", "Ts", "Te", "density")
GRIDGEN_AXES = (("vpar", ("vparGL", "vparGH", "vparGC", "dvpar")),
                ("vperp", ("vperpGL", "vperpGH", "vperpGC", "dvperp")),
                ("vperp1", ("vperp1GL", "vperp1GH", "vperp1GC", "dvperp1")),
                ("vperp2", ("vperp2GL", "vperp2GH", "vperp2GC", "dvperp2")),
                ("vp", ("vpGL", "vpGH", "vpGC", "dvp")),
                ("vq", ("vqGL", "vqGH", "vqGC", "dvq")),
                ("vphi", ("vphiGL", "vphiGH", "vphiGC", "dvphi")))


class KaosGridgenIn(C.Structure):
    _fields_ = ([("nq_ub", C.c_int64)]
                + [(n, C.c_double) for n in ("qGA", "qGB", "L", "phi", "ms", "ti", "te", "ns0", "nsnormfac")]
                + [(n, C.c_int32) for n in ("vperp3d", "symvpar", "cx", "device")]
                + [("nvpar_gp", C.c_int64), ("nvperp_gp", C.c_int64)]
                + [(n, C.c_double) for n in ("vpar_sigma_fac", "vperp_sigma_fac", "t_neut", "zns0_neut",
                                             "ns0_neut")])


class KaosGridgenDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nq", "nvpar", "nvperp", "nvperp1", "nvperp2", "nvp", "nvq", "nvphi")]


class KaosGridgenOut(C.Structure):
    _fields_ = ([(n, _dp) for n in GRIDGEN_CFG]
                + [("counts", C.POINTER(C.c_int64)), ("ns_neut", _dp),
                   ("nsnorm_clb", C.c_double), ("nsnorm_cub", C.c_double)]
                + [(n, _dp) for _, names in GRIDGEN_AXES[:2] for n in names] + [("d3v", _dp)]
                + [(n, _dp) for _, names in GRIDGEN_AXES[2:4] for n in names] + [("d3v3", _dp)]
                + [(n, _dp) for _, names in GRIDGEN_AXES[4:] for n in names]
                + [("ena_hh", _dp), ("ena_d33v", _dp)])


def load(path=LIB_PATH):
    """Load libkaos_hip.so and declare the signatures of every export."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(path)
    P = C.c_void_p
    i64p = C.POINTER(C.c_int64)
    i32p = C.POINTER(C.c_int32)
    lib.kaos_init.argtypes = [C.POINTER(KaosCfg), C.POINTER(P)]
    lib.kaos_set_grid.argtypes = [P, C.c_int32, C.POINTER(KaosGrid), C.c_int64]
    lib.kaos_upload_particles.argtypes = [P, C.c_int32, C.c_int64, _dp, _dp, _dp]
    lib.kaos_upload_vperp.argtypes = [P, C.c_int32, C.c_int64, _dp, _dp]
    lib.kaos_cell_counts.argtypes = [C.c_int32, _dp, _dp, C.c_double, i64p]
    lib.kaos_seed_particles.argtypes = [P, C.c_int32, i64p, _dp, C.c_uint64]
    lib.kaos_gridgen_dims_of.argtypes = [C.POINTER(KaosGridgenIn), C.POINTER(KaosGridgenDims)]
    lib.kaos_generate_grid.argtypes = [C.POINTER(KaosGridgenIn), C.POINTER(KaosGridgenOut)]
    lib.kaos_step.argtypes = [P, C.c_int64, _dp, C.c_int64, C.POINTER(KaosStepOut)]
    lib.kaos_step_noalloc.argtypes = [P, C.c_int64, _dp, C.c_int64, C.POINTER(KaosStepOut)]
    lib.kaos_run.argtypes = [P, C.c_int64, C.c_int64, i64p]
    lib.kaos_download_master.argtypes = [P, C.c_int32, i64p, i64p, _dp, _dp, _dp, _dp, _dp]
    lib.kaos_export_master.argtypes = [P, C.c_int32, C.POINTER(KaosExportInfo), i64p]
    lib.kaos_export_spinup.argtypes = [P, C.c_int32, C.POINTER(KaosSpinupInfo)]
    lib.kaos_download_master3d.argtypes = [P, C.c_int32, i64p, _dp]
    lib.kaos_read_spinup.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int64, C.POINTER(KaosSpinupState)] + [_dp] * 5
    lib.kaos_set_field.argtypes = [P, C.c_int32, _dp, C.c_int32]
    lib.kaos_count.argtypes = [P, C.c_int32, i64p]
    lib.kaos_download_cx.argtypes = [P, C.c_int32, _dp, _dp, _dp]
    lib.kaos_download_enas.argtypes = [P, C.c_int32, C.c_int64, i64p] + [_dp] * 6 + [i64p] * 2
    lib.kaos_download_ena_master.argtypes = [P, C.c_int32, i64p, _dp]
    lib.kaos_download_ena_moments.argtypes = [P, C.c_int32, i64p, _dp]
    lib.kaos_download_particles.argtypes = [P, C.c_int32, C.c_int64, i64p] + [_dp] * 10 + [i64p] * 3
    lib.kaos_download_vperp.argtypes = [P, C.c_int32, C.c_int64, i64p, _dp, _dp]
    lib.kaos_upload_state.argtypes = [P, C.c_int32, C.c_int64] + [_dp] * 12 + [i64p] * 3
    lib.kaos_upload_moments.argtypes = [P, C.c_int32, _dp]
    lib.kaos_resume.argtypes = [P, C.c_int64]
    lib.kaos_save_state.argtypes = [P, C.c_char_p]
    lib.kaos_load_state.argtypes = [P, C.c_char_p]
    lib.kaos_set_launch.argtypes = [P, C.c_int32, C.c_int32, C.c_int32]
    lib.kaos_get_launch.argtypes = [P, i32p, i32p, i32p, i32p]
    lib.kaos_set_launch_block.argtypes = [P, C.c_int32]
    lib.kaos_get_launch_block.argtypes = [P, i32p]
    lib.kaos_profile.argtypes = [P, C.c_int32]
    lib.kaos_profile_read.argtypes = [P, _dp, i64p, _dp]
    lib.kaos_last_error.argtypes = [P]
    lib.kaos_last_error.restype = C.c_char_p
    lib.kaos_finalize.argtypes = [P]
    for n in EXPORTS:
        if n != "kaos_last_error":
            getattr(lib, n).restype = C.c_int
    _lib = lib
    return lib


def ptr(a, ct=C.c_double):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))
