/*
 * kaos_hip.h -- C-ABI of libkaos_hip.so, the MI355X (gfx950) implementation
 * of the KAOS per-timestep kinetic hot path.
 *
 * The reference has no FFI: its hot path is a set of argument-less Fortran
 * module subroutines called from KineticSolverSub (_8_KineticSolver.f90:35-263)
 * that communicate through module globals (_KineticMainParams.f90:33-554).
 * These entry points replace those subroutine calls one for one, take plain
 * pointers and sizes (host memory, valid only during the call) and return an
 * int status (0 = ok).  A Fortran driver binds them with iso_c_binding.
 *
 * Index conventions follow the reference: config cells are 1..nq with
 * 0 = escaped through the lower boundary and -1 = through the upper one;
 * velocity cells are 1..n with 0 = none.  Per-cell arrays are passed as
 * contiguous 0-based C arrays (cell Q of the reference at [Q-1]); the 2-D
 * phase-space arrays are [Q-1][Vperpind-1][Vparind-1].
 */
#ifndef KAOS_HIP_H
#define KAOS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define KAOS_OK            0
#define KAOS_E_ARG         1   /* invalid argument or configuration */
#define KAOS_E_CAPACITY    2   /* particle or ENA capacity of a tube exceeded */
#define KAOS_E_DEVICE      3   /* HIP runtime error */
#define KAOS_E_STATE       4   /* call out of order (e.g. step before upload) */
#define KAOS_E_NAN         5   /* a NaN was produced in the particle state */
#define KAOS_E_IO          6   /* a file could not be written */

/* Precision forms. */
#define KAOS_FP64   0   /* every stage in binary64 */
#define KAOS_MIXED  1   /* binary32 state/RK4 stages; binary64 root solve, moment sums, gradient */

/* Run-wide configuration (replaces the flag copies of _1_SimParameterization.f90
 * and the allocation of DataTypeAllocBSub, _8_0_5_DataTypeAllocB.f90). */
typedef struct kaos_cfg {
  int32_t ntubes;        /* flux tubes (SpecieT(s)%NfT) */
  int32_t precision;     /* KAOS_FP64 or KAOS_MIXED */
  int32_t ndatfac;       /* master-step cadence ndatfacGT (>= 1) */
  int32_t device;        /* HIP device ordinal */
  uint64_t seed;         /* seed of the injection stream (used when no uniforms are passed) */
  int32_t reorder_every; /* master steps between on-device compaction + sort of the ions by
                            cell (0 = library default, 128; < 0 = never); results do not
                            depend on it */
  int32_t flags;         /* KAOS_FLAG_* (0 = defaults: autotuned launches, graph replay in kaos_run) */
} kaos_cfg;

#define KAOS_FLAG_NO_GRAPH     1   /* kaos_run enqueues every step eagerly instead of replaying a
                                      captured hipGraph of the step's kernels */
#define KAOS_FLAG_NO_AUTOTUNE  2   /* keep the default launch shape instead of timing candidates
                                      on the first master steps */
#define KAOS_FLAG_VPERP3D      4   /* ION2VPERPflag = 1: count on the 3-D (Vperp1, Vperp2, Vpar)
                                      grid of every tube and compute the 3-D moments
                                      (kaos_download_master3d) instead of the 2-D (Vperp, Vpar) ones */

/* One flux tube: grid, species and flags.  Replaces the FluxTubeT/QCellT/VCellT
 * grid fields (_KineticMainParams.f90:397-520) and the per-tube flags. */
typedef struct kaos_grid {
  int32_t nq;            /* config cells NqUB-NqLB+1 (>= 11: moment filter window) */
  int32_t nvpar;         /* NVparGT  (<= 255) */
  int32_t nvperp;        /* NVperpGT (<= 255) */
  int32_t grav;          /* GRAVflagT */
  int32_t eamb;          /* EAMBflagT */
  int32_t eambsign;      /* EAMBSIGNflagT */
  int32_t filter_pt;     /* M0MAfilterPt (odd) */
  int32_t lbcond;        /* LBCONDITIONflagT */
  int32_t ubcond;        /* UBCONDITIONflagT */
  int32_t filter_pt_m1par; /* M1ParMAfilterPt (odd; 0 = filter_pt): window of the filtered
                            parallel drift M1ParFiltAvrgRT (_8_7_AmbipolarEfield.f90:64-80) */
  int64_t n_inj_lb;      /* ions injected at NqLB per master step: nint(nsnormCLBGT/ranksize) */
  int64_t n_inj_ub;      /* same at NqUB */
  double p;              /* L-shell pGCGT */
  double cphi, sphi;     /* cos/sin of the tube longitude phiGCGT */
  double h;              /* time step hT [s] */
  double ms, qs;         /* species mass [kg] and charge [C] */
  double nsnormfac;      /* macro-particle normalisation nsnormfacT */
  double Ts_lb, Ts_ub;   /* ion temperature TsGT at NqLB / NqUB [K] */
  /* config grid, nq each (qGLGT, qGHGT, qGCGT, hqCGT, dqCGT, d3xCGT,
   * rGCGT, cos(thetaGCGT), ellGCGT, TeNT); edges must be contiguous:
   * qGL[Q+1] == qGH[Q]. */
  const double *qGL, *qGH, *qGC, *hqC, *dqC, *d3xC, *rGC, *cthGC, *ellGC, *Te;
  /* velocity grid (VparGLT/GHT/GCGT/dVparGT, nvpar; VperpGLT/GHT/GCGT/dVperpGT,
   * nvperp; d3vCT, nvperp*nvpar as [Vperpind-1][Vparind-1]); contiguous edges. */
  const double *vparGL, *vparGH, *vparGC, *dvpar;
  const double *vperpGL, *vperpGH, *vperpGC, *dvperp;
  const double *d3v;
  /* ICR wave heating (ICRflagT = 1; stochastic, ICRCOHERENCEflag = 0), with
   * the parameters of _ModelInput.f90:214-229, uniform over the cells as
   * SimParameterizationSub sets them (_1_SimParameterization.f90:1430-1450).
   * A zero-initialised struct has heating off. */
  int32_t icr;
  int32_t pad1;
  double icr_lambda_perp;  /* lambdaPerpp [m], 0 = infinite-wavelength limit */
  double icr_eta_lh;       /* EtaLHp */
  double icr_xi_perp1;     /* XiPerp1p */
  double icr_xi_perp2;     /* XiPerp2p */
  double icr_s0;           /* S0p [(V^2/m^2)/Hz] */
  double icr_f0;           /* f0p [Hz] (OmegaG0p = 2 pi f0p) */
  double icr_chi_perp1;    /* ChiPerp1p */
  double icr_chi_perp2;    /* ChiPerp2p */
  /* 3-D velocity grid (ION2VPERPflag = 1; used when the context has
   * KAOS_FLAG_VPERP3D): Vperp1 cells (Vperp1GLT/GHT/GCT/dVperp1GT, nvperp1),
   * Vperp2 cells (nvperp2), contiguous edges, and d3v3 = dV1 dV2 dVpar
   * (d3vCTp) as [Vperp1ind-1][Vperp2ind-1][Vparind-1], Vpar fastest
   * (_0_0_VelGridGenerator.f90:384-398).  The Vpar cells are the ones above. */
  int32_t nvperp1, nvperp2;
  const double *vperp1GL, *vperp1GH, *vperp1GC, *dvperp1;
  const double *vperp2GL, *vperp2GH, *vperp2GC, *dvperp2;
  const double *d3v3;
  /* Ion-neutral charge exchange (QEXCHANGEflagT = 1; needs KAOS_FLAG_VPERP3D,
   * whose M1Perp1/M1Perp2 moments the ion mean velocity reads).  At step 1
   * and every master step each config cell gets a Poisson collision count
   * (IonNeutralCollisionFrequencyA/BSub, _8_6_6_1:93-300, _8_6_6_2:35-77);
   * at every step of the next master interval the first nint(count) ions of
   * the cell (id order) become ENAs (_8_KineticSolver.f90:95-166,
   * IonNeutralCollisionsSub), which then move ballistically under gravity,
   * are counted per config cell and escape at LB/UB. */
  int32_t cx;              /* 1: charge exchange on in this tube */
  int32_t cx_filter;       /* MOMENTFILTERflagT: 1 = the filtered M1Par (window
                              filter_pt_m1par), M1Perp1, M1Perp2, M2Par (cx_filter_pt) */
  int32_t cx_filter_pt;    /* M1Perp1/M1Perp2/M2ParMAfilterPt (odd) */
  int32_t cx_pad;
  double t_neut;           /* TNeut [K] */
  double m_neut;           /* mNeut [kg] */
  double cx_sigma_scale;   /* factor on sigmaIonNeutRT; 1 = the reference's expression as written */
  const double* ns_neut;   /* nsnormCNeutGT(1, Q): neutral macro-particles per cell (nq) */
  int64_t ena_cap;         /* ENA slots of the tube (conversions over the whole run) */
  /* The ENA velocity grid (_0_0_VelGridGenerator.f90:482-857): |Vp| cells
   * (nvp), Vq cells (nvq), Vphi cells (nvphi), each with contiguous edges,
   * centres and widths (<= 255 cells); per config cell and bin
   * [Q-1][ip][iq][iphi] (Vphi fastest; nq * nvp * nvq * nvphi values) the
   * metric product hVp hVq hVphi and d33vCT, which depend on the cell's
   * thetaGC (_0_0_VelGridGenerator.f90:820-857).  An ENA is binned by the
   * cells of its (|Vp|, Vq, Vphi) at its pre-push position (KUA:1156-1290;
   * off the grid a component keeps its previous cell). */
  int32_t nvp, nvq, nvphi, ena_pad;
  const double *vpGL, *vpGH, *vpGC, *dvp;
  const double *vqGL, *vqGH, *vqGC, *dvq;
  const double *vphiGL, *vphiGH, *vphiGC, *dvphi;
  const double *ena_hh, *ena_d33v;
} kaos_grid;

/* Per-step summary. */
typedef struct kaos_step_out {
  int64_t n_alive;       /* live ions after the step, all tubes: the ions pushed this step
                            plus the ions injected (kaos_count summed over the tubes) */
  int64_t n_injected;    /* ions injected this step */
  int32_t master;        /* 1 when this step was a master step */
  int32_t status;        /* KAOS_OK or an error code raised on the device */
} kaos_step_out;

typedef struct kaos_ctx kaos_ctx;

/* Create a context on cfg->device.  Replaces DataTypeAllocBSub
 * (_8_0_5_DataTypeAllocB.f90:23) and the flag copies of SimParameterizationSub. */
int kaos_init(const kaos_cfg* cfg, kaos_ctx** out);

/* Set tube `tube`'s grid, species and flags and reserve `capacity` particle
 * slots for it.  Replaces the grid arrays built by SimParameterizationSub
 * (_0_ConfigGridGenerator.f90, _0_0_VelGridGenerator.f90) as read by
 * KineticSolverSub; GravfieldSub's cell values (_8_7_2_Gravfield.f90:57-60)
 * are computed here once. */
int kaos_set_grid(kaos_ctx* ctx, int32_t tube, const kaos_grid* g, int64_t capacity);

/* Initial ions of a tube: dipole q0 (q0T), parallel velocity Vpar0T and
 * perpendicular speed Vperp0T.  Positions follow from the root solve on the
 * tube's L-shell and velocities are Vpar0 b-hat.  Replaces x0T..Vz0T of
 * InitialConditionsSub (_7_InitialConditions.f90:51-101). */
int kaos_upload_particles(kaos_ctx* ctx, int32_t tube, int64_t n, const double* q0,
                          const double* vpar0, const double* vperp0);

/* Both perpendicular components (Vperp10T, Vperp20T; KSA:389-391) of the n
 * ions just uploaded to `tube`, before its first step; the Vperp cells follow
 * from |(vperp1, vperp2)|.  Without this call an uploaded ion has Vperp1 =
 * vperp0 and Vperp2 = 0 (only the magnitude enters unless ICR heating is on). */
int kaos_upload_vperp(kaos_ctx* ctx, int32_t tube, int64_t n, const double* vperp1, const double* vperp2);

/* Macro-ions per config cell of a tube: counts[Q] = (density[Q] /
 * nsnormfac) d3x[Q] truncated toward zero (the reference stores it in the
 * integer array nsnormC0, _KineticMainParams.f90:64); density in m^-3, d3x
 * the cell volumes.  Replaces nsnormC0 (_0_1_1_ConfigGridGeneratorB.f90:72-85)
 * and NsFARRp (_2_1_DensityProfileA1.f90:45) for one rank.  Host arithmetic
 * only: no context, no device.  KAOS_E_ARG for nq < 1, nsnormfac <= 0 or a
 * count that is not finite or does not fit in int64. */
int kaos_cell_counts(int32_t nq, const double* density, const double* d3x, double nsnormfac, int64_t* counts);

/* Grid set-up of one tube from the model inputs, computed on the device
 * (one thread per cell or bin).  Replaces ConfigGridGeneratorSub
 * (_0_ConfigGridGenerator.f90:88-668: the q grid, the grid root solve, the
 * metric factors and cell volumes), ConfigGridGeneratorSubB
 * (_0_1_1_ConfigGridGeneratorB.f90:25-218: the hydrostatic ion density and
 * its macro-ion counts nsnormC0, the neutral density), VelGridGeneratorSub
 * (_0_0_VelGridGenerator.f90:25-857: the ion velocity grid, 2-D or 3-D, and
 * the ENA grid with its metric) and the re-index onto the loop's cells
 * (_1_SimParameterization.f90:1156-1179: loop cell Q is set-up cell Q + 1).
 * NqLB is 1 and the grid-point strides (dq, ddVpar, ...) are 1: other values
 * index past the reference's arrays. */
typedef struct kaos_gridgen_in {
  int64_t nq_ub;           /* NqUB (NqLB = 1): nq = NqUB config cells */
  double qGA, qGB;         /* q of the first and the last grid point */
  double L, phi;           /* Lshell, phiLshell (phi: carried to kaos_grid only) */
  double ms;               /* species mass msT [kg] */
  double ti, te;           /* Ti(1), Te(1) [K]: TsPerp = TsPar = Ti, Te0 = Te */
  double ns0;              /* ion density at set-up cell Qindns0 = 1 [m^-3] */
  double nsnormfac;
  int32_t vperp3d;         /* ION2VPERPflag: 1 = the (Vperp1, Vperp2, V||) grid, 0 = (Vperp, V||) */
  int32_t symvpar;         /* SYMVPARflag: 1 = V|| grid symmetric about 0 */
  int32_t cx;              /* QEXCHANGEflag: the neutral density and the ENA grid */
  int32_t device;          /* HIP device ordinal */
  int64_t nvpar_gp;        /* NVparGpF (even, >= 4) */
  int64_t nvperp_gp;       /* NVperpGpF (even, >= 4) */
  double vpar_sigma_fac;   /* VparsigmaFac(1) */
  double vperp_sigma_fac;  /* VperpsigmaFac(1) (also the ENA Vp/Vphi factor) */
  double t_neut;           /* TNeut(1) [K] (cx) */
  double zns0_neut;        /* zns0Neut(1) [m] (cx) */
  double ns0_neut;         /* ns0Neut(1) [m^-3] (cx) */
} kaos_gridgen_in;

/* Cell counts of the grids kaos_generate_grid writes. */
typedef struct kaos_gridgen_dims {
  int32_t nq;              /* NqUB */
  int32_t nvpar;           /* NVparGpF - 2 */
  int32_t nvperp;          /* NVperpGpF - 2 (2-D grid), else 0 */
  int32_t nvperp1, nvperp2;/* NVperpGpF - 2 each (3-D grid), else 0 */
  int32_t nvp, nvq, nvphi; /* ENA grid (cx): NVperpGpF - 2, NVparGpF - 2, NVperpGpF - 2, else 0 */
} kaos_gridgen_dims;

/* Caller-allocated outputs (any pointer may be NULL).  Per config cell
 * (nq): the kaos_grid arrays of the tube, the ion temperature Ts0 = Ti/3 +
 * 2 Ti/3 and Te0, the hydrostatic density nsC [m^-3], its macro-ion counts
 * nsnormC0 (truncated: an integer array in the reference) and the neutral
 * macro-particle counts nsnormCNeutGT (truncated likewise).  The velocity
 * grids as kaos_grid takes them (V|| cells with the centre 0 replaced by
 * 1e-15, _0_0_VelGridGenerator.f90:1000, 1168); 2-D: d3v = 2 pi |Vperp_c|
 * dVperp dV|| as [Vperp][V||]; 3-D: d3v3 = dV1 dV2 dV|| as [V1][V2][V||].
 * ENA (cx): the Vp, Vq, Vphi cells and per config cell and bin
 * [Q][ip][iq][iphi] hh = hVp hVq hVphi and d33vCT, from the theta of set-up
 * cell Q (the ENA grid is built before the re-index, so loop cell Q takes
 * the centre one cell below its own). */
typedef struct kaos_gridgen_out {
  double *qGL, *qGH, *qGC, *hqC, *dqC, *d3xC, *rGC, *cthGC, *ellGC, *Ts, *Te;
  double *density;         /* nsC */
  int64_t *counts;         /* nsnormC0 of the loop's cells */
  double *ns_neut;         /* nsnormCNeutGT (cx) */
  double nsnorm_clb;       /* nsnormCLBG = nsnormC0 of loop cell 1 (the LB injection per
                              master step is nint(nsnorm_clb / ranks)) */
  double nsnorm_cub;       /* nsnormCUBG = nsnormC0 of loop cell nq */
  double *vparGL, *vparGH, *vparGC, *dvpar;
  double *vperpGL, *vperpGH, *vperpGC, *dvperp, *d3v;            /* 2-D */
  double *vperp1GL, *vperp1GH, *vperp1GC, *dvperp1;               /* 3-D */
  double *vperp2GL, *vperp2GH, *vperp2GC, *dvperp2, *d3v3;
  double *vpGL, *vpGH, *vpGC, *dvp;                               /* ENA */
  double *vqGL, *vqGH, *vqGC, *dvq;
  double *vphiGL, *vphiGH, *vphiGC, *dvphi;
  double *ena_hh, *ena_d33v;
} kaos_gridgen_out;

/* The grid sizes of `in` (host arithmetic, no device).  KAOS_E_ARG for
 * NqUB < 11, odd or < 4 grid-point counts, a cell count above 255, or
 * non-positive masses, temperatures or normalisation. */
int kaos_gridgen_dims_of(const kaos_gridgen_in* in, kaos_gridgen_dims* dims);

/* Generate the grids of `in` on device in->device into `out`. */
int kaos_generate_grid(const kaos_gridgen_in* in, kaos_gridgen_out* out);

/* Initial ions of a tube generated on the device, in place of
 * kaos_upload_particles + kaos_upload_vperp: counts[Q] ions at the centre
 * qGC of config cell Q (DensityProfileBSub, _4_DensityProfileB.f90:40-46),
 * signed radius from the root solve (DipolePolynomialSolverSub), and a
 * Maxwell-Boltzmann velocity of temperature temperature[Q] (K) projected on
 * the field line at that point (VelocityDistributionSub,
 * _6_VelocityDistribution.f90:60-85; MBVelocityDistributionSub).  Ions are
 * numbered cell by cell (InitialConditionsSub, _7_InitialConditions.f90:72-87);
 * ion i draws its six uniforms from Philox4x64-10 keyed by seed (counter i,
 * tube).  Replaces the reference's GaussianRNGSub draws, which come from
 * Fortran's random_number and cannot be reproduced. */
int kaos_seed_particles(kaos_ctx* ctx, int32_t tube, const int64_t* counts, const double* temperature,
                        uint64_t seed);

/* Advance every tube by time step n (n = 1 first).  One call replaces one
 * pass of ComputationalTimeStepLoop (_8_KineticSolver.f90:54-239):
 * BoundaryConditionsSub, AmbipolarEfieldSub, GravfieldSub, KineticSolverA/BSub
 * (KineticRK4UpdateSub push loop and LB/UB injection), ParticleCountsSub,
 * DistributionFunctionsSub and the six moment subroutines.
 * uniforms: NULL for the seeded device stream, else 6 uniforms in (0,1] per
 * injected ion of this step, tube-major, LB ions before UB ions
 * (the GaussianRNGSub draws of KSB:326/623); n_uniforms their count.
 * out may be NULL; when given the call waits for the step to finish. */
int kaos_step(kaos_ctx* ctx, int64_t n, const double* uniforms, int64_t n_uniforms,
              kaos_step_out* out);

/* kaos_step for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); it has
 * no counterpart in the reference.  Same arguments and results as kaos_step,
 * and it calls no hipMalloc/hipFree or other allocating routine: the caller's
 * uniforms are staged in a fixed device array of the library (at most
 * 6 * 65536 doubles: up to 65536 injected ions per step, else KAOS_E_ARG),
 * the step is enqueued without graph replay, and the autotuner and kaos_profile
 * timing skip it (the current launch shape is used).  It returns after the step
 * has finished.  The staging array is one per GPU, shared by every context on
 * it: calls that overlap on one GPU race. */
int kaos_step_noalloc(kaos_ctx* ctx, int64_t n, const double* uniforms, int64_t n_uniforms,
                      kaos_step_out* out);

/* Steps n_first .. n_first+nsteps-1 with the seeded stream and no host
 * synchronisation between steps; returns after the last one finishes.
 * particle_steps (may be NULL) receives the ions pushed summed over steps. */
int kaos_run(kaos_ctx* ctx, int64_t n_first, int64_t nsteps, int64_t* particle_steps);

/* Outputs of the last master step of a tube.  Any pointer may be NULL.
 * nq_counts (nq): NqReNormT; nph_counts (nq*nvperp*nvpar): NphReNormRT;
 * NqRT (nq); M (6*nq): M0phRT, M1PerpphRT, M1ParphRT, M2phRT, M2PerpphRT,
 * M2ParphRT; EA (nq): EAmagRT; EG (nq): EGmagRT; esc (4): LB and UB escaped
 * ion counts and the LB/UB escape fluxes LBoutfluxIonRT/UBoutfluxIonRT
 * (_8_0_7_BoundaryConditions.f90:201-208, 278-285).  With KAOS_FLAG_VPERP3D
 * the 2-D counts are zero and M holds the moments both branches define (M0,
 * M1Par, M2, M2Par; M1Perp and M2Perp zero): the 3-D ones come from
 * kaos_download_master3d. */
int kaos_download_master(kaos_ctx* ctx, int32_t tube, int64_t* nq_counts, int64_t* nph_counts,
                         double* NqRT, double* M, double* EA, double* EG, double* esc);

/* Spin-up (restart) output of DataExport2Sub (SPINUPflag = 1,
 * _8_0_2_DataExport2.f90:704-802): the caller's grid bounds, injection
 * normalisations and replenish flags, which this path does not hold. */
typedef struct kaos_spinup_info {
  const char* dir;         /* Densitydatadir, NUL-terminated (trailing '/' optional) */
  int32_t species;         /* s (1-based) */
  int32_t flux_tube;       /* f (1-based) */
  int64_t nq_lb, nq_ub;    /* NqLBT(1), NqUBT(1) */
  double nsnorm_clb;       /* nsnormCLBGT(1) */
  double nsnorm_cub;       /* nsnormCUBGT(1) */
  int64_t lb_replenish;    /* LBREPLENISHflagT(1) */
  int64_t ub_replenish;    /* UBREPLENISHflagT(1) */
} kaos_spinup_info;

/* The spin-up files of the last master step (step 1 or a master step) of a
 * tube, <s>_<f>_<Name>fort.bin in info->dir, replaced at every call as the
 * reference replaces them at every export: NqLBT, NqUBT (int64),
 * DensityOutputRT = M0phRT, TemperatureOutputRT = (2/3) M2ParphRT/kB +
 * (2/3) M2Perp1phRT/kB + (2/3) M2Perp2phRT/kB (_8_6_6_SecondParIonMoment.f90:
 * 659-673), EAInertialOutputRT (0: EAINERTIALflag off), EAPressureOutputRT,
 * EAmagOutputRT (_8_7_AmbipolarEfield.f90:1419-1421), nsnormCLBOutput,
 * nsnormCUBOutput (real64), LBREPLENISHflagTOutput, UBREPLENISHflagTOutput
 * (int64).  Needs KAOS_FLAG_VPERP3D: the reference allocates M2Perp1phRT and
 * M2Perp2phRT only with ION2VPERPflag = 1 (_8_0_5_DataTypeAllocB.f90:507-519),
 * so its spin-up temperature is defined only there (KAOS_E_STATE otherwise).
 * KAOS_E_IO if a file cannot be written. */
int kaos_export_spinup(kaos_ctx* ctx, int32_t tube, const kaos_spinup_info* info);

/* The scalars of a previous run's spin-up files (SPINUPflag = 0 restart,
 * _1_SimParameterization.f90:801-946). */
typedef struct kaos_spinup_state {
  int64_t nq_lb, nq_ub;    /* NqLBp, NqUBp (NqLBTfort.bin, NqUBTfort.bin) */
  double nsnorm_clb;       /* nsnormCLBInput: the count of the lower ghost cell
                              (nsnormC0(NqLB), _0_1_1_ConfigGridGeneratorB.f90:76-78).
                              The LB injection of the restarted run is not this value
                              but nsnormCLBGT = nsnormC0(NqLB + 1) (:189), the first
                              computational cell's count kaos_cell_counts gives:
                              nint(counts[0] / ranks) ions per master step (BC:747-750) */
  double nsnorm_cub;       /* nsnormCUBInput */
  int64_t lb_replenish;    /* LBREPLENISHflagInput */
  int64_t ub_replenish;    /* UBREPLENISHflagInput */
  int64_t nq;              /* cells of the arrays: nq_ub - nq_lb + 1 */
} kaos_spinup_state;

/* Read the spin-up files <s>_<f>_<Name>fort.bin of Densitydatadir `dir` as the
 * reference's restart does (_1_SimParameterization.f90:801-946: Fortran
 * stream-unformatted, little-endian int64/binary64, no record markers):
 * NqLBT, NqUBT, nsnormCLBOutput, nsnormCUBOutput, DensityOutputRT,
 * TemperatureOutputRT, EAInertialOutputRT, EAPressureOutputRT,
 * EAmagOutputRT, LBREPLENISHflagTOutput, UBREPLENISHflagTOutput.  The arrays
 * (nq values each, any pointer may be NULL) are the reference's
 * DensityInputT (m^-3), TemperatureInputT (K) and the three EA inputs
 * (V/m); a restart seeds a tube from them: kaos_cell_counts(density) for
 * nsnormC0 (_0_1_1_ConfigGridGeneratorB.f90:75-86), kaos_seed_particles
 * with the temperatures (_6_VelocityDistribution.f90:70-73) and
 * kaos_set_field with ea_mag (_8_7_AmbipolarEfield.f90:1375-1407).  Host
 * file reads only: no context, no device.  KAOS_E_IO if a file is missing
 * or not of the size its name implies (nq from NqLBT/NqUBT), KAOS_E_ARG if
 * nq exceeds nq_cap or nq < 1. */
int kaos_read_spinup(const char* dir, int32_t species, int32_t flux_tube, int64_t nq_cap,
                     kaos_spinup_state* st, double* density, double* temperature, double* ea_inertial,
                     double* ea_pressure, double* ea_mag);

/* Source of a tube's ambipolar field EAmagRT at master steps. */
#define KAOS_FIELD_COMPUTED    0  /* AmbipolarEfieldSub's stencils (default) */
#define KAOS_FIELD_INPUT_FIRST 1  /* the given field at step 1, computed later
                                     (SPINUPflag = 0 with EAMBSELFCONSISTflag = 1
                                     or dNTe /= 0, _8_7_AmbipolarEfield.f90:1375-1390) */
#define KAOS_FIELD_INPUT_FIXED 2  /* the given field at step 1 and every master
                                     step (SPINUPflag = 0, EAMBSELFCONSISTflag = 0,
                                     dNTe = 0, _8_7_AmbipolarEfield.f90:1392-1407) */

/* Set the input field of a tube (nq values of EAmagInputT, V/m; may be NULL
 * with KAOS_FIELD_COMPUTED) and its mode.  The input replaces the computed
 * EAmagRT after the stencils and their LB override, as the reference's
 * assignment follows them; the per-ion interpolation and the injected ions'
 * AEAmag then read it.  Allowed after kaos_set_grid and between steps; it
 * takes effect at the next field evaluation (step 1 or a master step). */
int kaos_set_field(kaos_ctx* ctx, int32_t tube, const double* EA, int32_t mode);

/* The 3-D outputs of the last master step of a tube (KAOS_FLAG_VPERP3D):
 * n3_counts (nq*nvperp1*nvperp2*nvpar, [Q][V1][V2][Vpar], Vpar fastest):
 * N2PerpphReNormRT (_8_3_1_IonParticleCounts.f90:336-361); M (8*nq): M0phRT,
 * M1Perp1phRT, M1Perp2phRT, M1ParphRT, M2phRT, M2Perp1phRT, M2Perp2phRT,
 * M2ParphRT (3-D branches of _8_6_1.._8_6_6).  Either pointer may be NULL.
 * In this mode kaos_download_master's M holds M0, 0, M1Par, M2, 0, M2Par
 * (M1Perp and M2Perp are 2-D quantities) and its nph_counts are zero. */
int kaos_download_master3d(kaos_ctx* ctx, int32_t tube, int64_t* n3_counts, double* M);

/* What the export files hold that the hot path does not compute. */
typedef struct kaos_export_info {
  const char* dir;       /* dataexportdir; a '/' is put between it and the names if missing */
  int32_t species;       /* s of the file names (1-based, SpecieT index) */
  int32_t flux_tube;     /* f of the file names (1-based, FluxTubeT index) */
  double ns0;            /* ns0(1), reference density [m^-3]: ns0GTfort.bin */
  double phi;            /* phiGCGT of the tube's cells [rad]: phiGCGTfort.bin */
  const double* theta;   /* thetaGCGT(nn, :), nq values [rad]: thetaGCGTfort.bin */
  const double* Ts;      /* TsGT(nn, :), nq values [K]: TsGTfort.bin */
} kaos_export_info;

/* The files DataExport2Sub writes for one tube at master step nn
 * (_8_0_2_DataExport2.f90:25-704), from the last step, which must be step 1
 * (nn = 1) or a master step n (nn = n/ndatfac + 1; KSB:1090-1097, 1214-1221,
 * KSA:419-428).  Names "<nn>_<s>_<f>_<Name>fort.bin" and, per config cell Q,
 * "<nn>_<s>_<f>_<Q>_<Name>fort.bin"; each file is a Fortran stream-unformatted
 * write: the values in array order, little-endian binary64 (ndatfacGT:
 * int64), no record markers.  The flag set is that of this path
 * (FLUIDIONEXPORT, PHASEIONDISTRIB, the moment and filter flags on;
 * ION2VPERP, QEXCHANGE, SPINUP, EAINERTIAL, EPAR off): TimeT, ndatfacGT,
 * ns0GT, qGCGT, pGCGT, rGCGT, phiGCGT, thetaGCGT, d3xCGT, TsGT; NphRT per cell
 * ((Vperp, Vpar) column-major: NphReNormRT nsnormfac/d3x, 0 where the count
 * is 0); M0phRT, M1PerpphRT, M1ParphRT, M2phRT, M2PerpphRT, M2ParphRT;
 * M0FiltAvrgRT, M1ParFiltAvrgRT and M1Perp1/M1Perp2/M2ParFiltAvrgRT (zeros:
 * only charge exchange sets them); EAInertialRT (zeros), EAPressureRT,
 * EAmagRT, EGmagRT; EPmagRT and PhiParRT (zeros).  With KAOS_FLAG_VPERP3D
 * (ION2VPERP on, _8_0_2_DataExport2.f90:123-161, 276-307, 446-490) the cell
 * files are N2PerpphRT ((V1, V2, Vpar) column-major), TeNT and d3vCTp instead
 * of NphRT, and M1Perp1/M1Perp2phRT, M2Perp1/M2Perp2phRT replace M1PerpphRT,
 * M2PerpphRT.  nn_out (may be NULL) receives nn.  KAOS_E_IO if a file cannot
 * be written. */
int kaos_export_master(kaos_ctx* ctx, int32_t tube, const kaos_export_info* info, int64_t* nn_out);

/* Charge-exchange state of a tube's last master step (kaos_grid.cx): per
 * config cell sigmaIonNeutRT, nuIonNeutRT (the capped Poisson count, -2 for
 * an empty cell or sigma = 0) and nuIonNeutPoiT (the one-rank share, ions
 * converted per step is its nint); any pointer may be NULL. */
int kaos_download_cx(kaos_ctx* ctx, int32_t tube, double* sigma, double* nu, double* poi);

/* The live ENAs of a tube in id order (the reference's array order: an ENA
 * keeps its ion's index): Cartesian position and velocity, id, and Qindk1 of
 * the last push (_8_1_1_KineticUpdateA.f90:221-273 on the ENA's q). */
int kaos_download_enas(kaos_ctx* ctx, int32_t tube, int64_t n_cap, int64_t* n_out, double* x, double* y,
                       double* z, double* Vx, double* Vy, double* Vz, int64_t* id, int64_t* qind);

/* ENA outputs of the last master step: nq_ena (nq): NqReNormENAT
 * (ENAParticleCountsSub, _8_3_2:33-60); esc (4): LB and UB escaped ENA
 * counts and the fluxes LBoutfluxENART/UBoutfluxENART
 * (_8_0_7_BoundaryConditions.f90:309-470).  Either pointer may be NULL. */
int kaos_download_ena_master(kaos_ctx* ctx, int32_t tube, int64_t* nq_ena, double* esc);

/* ENA phase-space counts NphReNormENAT of the last master step, [Q][ip][iq][iphi]
 * (nq * nvp * nvq * nvphi; ENAParticleCountsSub, _8_3_2:300-361), and M (8*nq):
 * M0phENART, M1PphENART, M1QphENART, M1PHIphENART, M2phENART, M2PphENART,
 * M2QphENART, M2PHIphENART of ENADistributionFunctionsSub and the ENA moment
 * subroutines (_8_4_2, _8_9, _8_10_1.._8_10_8; second moments in eV).  Either
 * pointer may be NULL. */
int kaos_download_ena_moments(kaos_ctx* ctx, int32_t tube, int64_t* counts, double* M);

/* Number of live ions of a tube. */
int kaos_count(kaos_ctx* ctx, int32_t tube, int64_t* n);

/* Live ions of a tube in the reference's order (survivors in order, then
 * injected ions), any pointer may be NULL; n_cap is the capacity of the
 * caller's arrays.  qk4: stage-4 dipole q of the last push at a step before a
 * master step; aea/ag: AEAmagN/AGmagN.  An ion not yet pushed since its
 * injection has qk4 = aea = ag = 0 (the reference leaves a new slot's values
 * unset until that push).  qind/vparind/vperpind: Qindk1,
 * Vparindk1, Vperpindk1 of the last push, integer(kind=dp) as the reference
 * declares them (_8_0_4_DataTypeAllocA.f90:35-83, dp = 8, _ModelInput.f90:15).
 * In the mixed form vparind is the cell of the stored Vpar (the index the next
 * push bins). */
int kaos_download_particles(kaos_ctx* ctx, int32_t tube, int64_t n_cap, int64_t* n_out,
                            double* x, double* y, double* z, double* Vx, double* Vy, double* Vz,
                            double* vperp, double* qk4, double* aea, double* ag,
                            int64_t* qind, int64_t* vparind, int64_t* vperpind);

/* Both perpendicular components (the reference's Vperp1, Vperp2 arrays) of
 * the live ions of a tube, in the order of kaos_download_particles; either
 * pointer may be NULL. */
int kaos_download_vperp(kaos_ctx* ctx, int32_t tube, int64_t n_cap, int64_t* n_out, double* vperp1,
                        double* vperp2);

/* Hand a tube the per-particle state the Fortran driver holds
 * (_8_0_4_DataTypeAllocA.f90:35-83): x, y, z, Vx, Vy, Vz, Vperp1, Vperp2,
 * Vperp, AEAmag, AGmag, qk4 (_8_1_3_KineticUpdateC.f90:31-62, the field
 * coordinate of the next master step) and the int64 cell indices Qindk1,
 * Vparindk1, Vperpindk1 of the last push; vperp1/vperp2 may be NULL (then
 * Vperp1 = Vperp, Vperp2 = 0).  The tube's ions become these n ions, numbered
 * in the given order.  Allowed before the first step (in place of
 * kaos_upload_particles) and between steps (the tube's other state stays).
 * The Cartesian state is read as KineticUpdateASub reads it: r = |(x, y, z)|
 * (_8_1_1_KineticUpdateA.f90:117), Vpar = b-hat . V (:978-980); so a
 * download followed by this upload continues the run to rounding (the
 * Cartesian triple of a line point does not fix its radius to the last bit;
 * kaos_save_state / kaos_load_state continue bit for bit).  KAOS_E_ARG for a
 * cell index out of range or a position at the origin or not finite. */
int kaos_upload_state(kaos_ctx* ctx, int32_t tube, int64_t n, const double* x, const double* y, const double* z,
                      const double* Vx, const double* Vy, const double* Vz, const double* vperp1,
                      const double* vperp2, const double* vperp, const double* aea, const double* ag,
                      const double* qk4, const int64_t* qind, const int64_t* vparind, const int64_t* vperpind);

/* The moments of the last master step of a tube (6*nq, the M of
 * kaos_download_master), which the next master step's moment filter and
 * ambipolar field read (AmbipolarEfieldSub uses M*phRT(nn-1),
 * _8_7_AmbipolarEfield.f90:40-82): for a context continuing another run. */
int kaos_upload_moments(kaos_ctx* ctx, int32_t tube, const double* M);

/* A context that has not stepped continues another run after its step n_last:
 * the next kaos_step is n_last + 1 (the seeded streams are keyed by the step).
 * Needs every tube's ions (kaos_upload_state) and, for the fields of the next
 * master step, kaos_upload_moments.  Replaces the loop counter of
 * ComputationalTimeStepLoop (_8_KineticSolver.f90:54). */
int kaos_resume(kaos_ctx* ctx, int64_t n_last);

/* Restart files of the hot path (no counterpart in the reference, whose only
 * restart is re-seeding from the spin-up density, _1_SimParameterization.f90:
 * 801-930): kaos_save_state writes the context's step state, per-cell outputs
 * and every tube's ion state as stored on the device to `path`;
 * kaos_load_state reads it into a context that has not stepped and was
 * created with the same kaos_cfg and grids (else KAOS_E_ARG), which then
 * continues bit for bit as the saved one would have.  KAOS_E_IO on a file
 * error. */
int kaos_save_state(kaos_ctx* ctx, const char* path);
int kaos_load_state(kaos_ctx* ctx, const char* path);

/* Launch shape of the push kernels: ions are processed in tiles of 1024, a
 * group of `tiles_per_group` tiles (1..63) shares one block-local LDS window of
 * `hist_words` 32-bit words (two 16-bit phase-space counters each; 0 = all
 * counts go to global atomics); `persistent` != 0 launches a few blocks per CU
 * that take groups from a counter, else one block per group.  Setting it ends
 * the autotuner.  Results do not depend on the launch shape.  No counterpart in
 * the reference (its loop is serial, KSB:139-1041). */
int kaos_set_launch(kaos_ctx* ctx, int32_t tiles_per_group, int32_t hist_words, int32_t persistent);
/* The launch shape in use and whether the autotuner is still timing candidates. */
int kaos_get_launch(kaos_ctx* ctx, int32_t* tiles_per_group, int32_t* hist_words, int32_t* persistent,
                    int32_t* tuning);
/* Threads per block of the push (256: four ions of a 1024-ion tile per
 * thread, 512: two), a dimension the autotuner times along with the shape
 * above; only the plain push (config cells <= 256, no charge exchange, and in
 * the mixed form no heating) has the 512-thread variant, the others run 256.
 * Setting it ends the autotuner.  Results do not depend on it.  No
 * counterpart in the reference. */
int kaos_set_launch_block(kaos_ctx* ctx, int32_t block);
int kaos_get_launch_block(kaos_ctx* ctx, int32_t* block);

/* Kernel timing: when enabled, hipEvents on the library's stream bracket every
 * push launch and every kaos_run call.  kaos_profile_read returns the summed
 * push-kernel time, the number of push launches and the summed kaos_run time
 * since the last read (milliseconds).  The reference's only timing is
 * cpu_time progress prints (_KineticMain.f90:49-52); this has no counterpart. */
int kaos_profile(kaos_ctx* ctx, int32_t enable);
int kaos_profile_read(kaos_ctx* ctx, double* push_ms, int64_t* push_launches, double* run_ms);

/* Text of the last error of a context (or of kaos_init when ctx is NULL). */
const char* kaos_last_error(kaos_ctx* ctx);

/* Release every device buffer of the context. */
int kaos_finalize(kaos_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* KAOS_HIP_H */
