/*
 * kaos_oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU restatement (FP64, one
 * thread, plain loops) of the KAOS per-timestep kinetic hot path, used by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker of the HIP path.  Nothing in the shipped library links or calls it.
 *
 * It follows the reference's call sequence and data layout:
 *   - per-ion state: the reference's Cartesian ion sits on its field line
 *     with V = V|| b-hat after every push (_8_1_KineticRK4Update.f90:170-314),
 *     so it is held as (signed radius rs, V||) and expanded to x,y,z,Vx,Vy,Vz
 *     on output; plus Vperp, the reference's per-ion fields AEAmag/AGmag
 *     (current) and AEAmagN/AGmagN (next), qk4 and the integer cell indices
 *     Qindk1/Vparindk1/Vperpindk1;
 *   - KineticSolverSub step order (_8_KineticSolver.f90:54-239):
 *     BoundaryConditionsSub -> [master] AmbipolarEfieldSub/GravfieldSub ->
 *     KineticSolverA/BSub (rotation, push loop, injection) ->
 *     ParticleCounts -> DistributionFunctions -> 6 moments;
 *   - linear cell scans (first match), stable pack for removal, brute-force
 *     equivalent histogram, column-major velocity sums.
 * The per-particle RK4/root-solve arithmetic is the shared scalar header
 * kaos_dipole.h; its agreement with the reference's own Fortran push
 * (compiled from the reference sources under oracle/ref) is tested separately.
 *
 * Parity pins for each function are listed in the test files; nothing here
 * is "parity unpinned" except where a test module says so.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "../kaos_amd/csrc/kaos_dipole.h"

typedef struct {
  /* scalars */
  int nq, nvpar, nvperp;
  int grav, eamb, eambsign, filter_pt;
  int ndatfac;
  int lbcond, ubcond;
  int64_t n_inj_lb, n_inj_ub;     /* nint(nsnormC{L,U}BGT/ranksize), BC:747-750, 806-809 */
  double p, cphi, sphi, h, ms, qs, nsnormfac;
  double Ts_lb, Ts_ub;            /* TsGT at NqLB / NqUB */
  int tube;                       /* tube number used to key the seeded stream */
  uint64_t seed;
  /* config grid (nq each), QCellT/FluxTubeT arrays */
  const double *qGL, *qGH, *qGC, *hqC, *dqC, *d3xC, *rGC, *cthGC, *ellGC, *Te;
  /* velocity grid */
  const double *vparGL, *vparGH, *vparGC, *dvpar;       /* nvpar  */
  const double *vperpGL, *vperpGH, *vperpGC, *dvperp;   /* nvperp */
  const double *d3v;                                    /* nvperp*nvpar, [iperp][ipar] */
  /* ICR wave heating (ICRflagT, stochastic; _ModelInput.f90:214-229) */
  int icr;
  double icr_lambda_perp, icr_eta_lh, icr_xi_perp1, icr_xi_perp2, icr_s0, icr_f0, icr_chi_perp1,
      icr_chi_perp2;
  int filter_pt_m1par;            /* M1ParMAfilterPt; 0 = filter_pt */
  /* 3-D velocity grid (ION2VPERPflag = 1): Vperp1, Vperp2 cells; d3v3 is
   * [i1][i2][ipar] (Vpar fastest), dV1 dV2 dVpar */
  int v3d, nv1, nv2;
  const double *v1GL, *v1GH, *v1GC, *dv1, *v2GL, *v2GH, *v2GC, *dv2, *d3v3;
  /* charge exchange (QEXCHANGEflagT = 1, 3-D grid): MOMENTFILTERflagT, the
   * M1Perp1/M1Perp2/M2Par filter window (M1Par: filter_pt_m1par), TNeut,
   * mNeut, nsnormCNeutGT per cell, ENA slots */
  int cx, cx_filter, cx_filter_pt, pad_cx;
  double t_neut, m_neut, cx_sigma_scale;
  const double *nsneut;
  int64_t ena_cap;
  /* ENA velocity grid: Vp, Vq, Vphi cells; per config cell and bin
   * [Q][ip][iq][iphi] the metric product hVp hVq hVphi and d33vCT */
  int nvp, nvq, nvphi, pad_ena;
  const double *vpGL, *vpGH, *vpGC, *dvp, *vqGL, *vqGH, *vqGC, *dvq, *vphiGL, *vphiGH, *vphiGC, *dvphi;
  const double *ena_hh, *ena_d33v;
} orc_cfg;

typedef struct {
  orc_cfg c;
  int64_t ns, cap;
  double *rs, *vs, *vperp, *aea, *ag, *aeaN, *agN, *qk4, *q0, *vpar0;
  double *v1, *v2;                 /* Vperp1, Vperp2 */
  double *ck4;                     /* cos(thetak4) of the last push (KUC), the frame of a conversion */
  int *qind, *vparind, *vperpind;
  int64_t *id;                     /* upload index, then injection order (keys the heating stream) */
  int64_t next_id;
  int master_last;                 /* last step was a master step */
  int64_t nn;                      /* index of the last master step */
  /* per master step outputs */
  double *nq_counts, *nph_counts;  /* NqReNormT, NphReNormRT (integer valued) */
  uint32_t *nq_u, *nph_u;
  double *NqRT, *M;                /* M: 6*nq: M0, M1perp, M1par, M2, M2perp, M2par */
  double *Mprev;                   /* moments of the previous master step (6*nq) */
  double *EA, *EG, *M0F;           /* EAmagRT(nn), EGmagRT(nn), M0FiltAvrgRT(nn) */
  double *M1F;                     /* M1ParFiltAvrgRT(nn) */
  int *i1ind, *i2ind, *ip3ind;     /* 3-D velocity cells (Vperp1indk1, Vperp2indk1, Vparindk1) */
  uint32_t *n3_u;                  /* N2PerpphReNormT, [Q][i1][i2][ipar] */
  double *M3;                      /* 8*nq: M0, M1Perp1, M1Perp2, M1Par, M2, M2Perp1, M2Perp2, M2Par */
  double esc_lb, esc_ub, flux_lb, flux_ub;
  int64_t injected;
  double *EAin;                    /* EAmagInputT of a restart (nq), or NULL */
  int ea_mode;                     /* 0 computed, 1 input at step 1, 2 input at every master step */
  /* charge exchange: per cell sigmaIonNeutRT, nuIonNeutRT, nuIonNeutPoiT of the
   * last master step; the ENAs (Cartesian state, id, Qindk1) and their
   * master-step counts NqReNormENAT and LB/UB escapes */
  double *cx_sigma, *cx_nu, *cx_poi;
  double *ex, *ey, *ez, *evx, *evy, *evz;
  int64_t *eid;
  int *eq, *evp, *evq, *evphi;     /* ENA Qindk1 and Vp/Vq/Vphi cells (stale off the grid) */
  uint32_t *ena_hist;              /* NphReNormENAT per cell, [Q][ip][iq][iphi] */
  double *ena_M;                   /* 8*nq: M0, M1P, M1Q, M1PHI, M2, M2P, M2Q, M2PHI of the ENAs */
  int64_t ens, ena_made;
  uint32_t *nq_ena;
  double ena_esc_lb, ena_esc_ub, ena_flux_lb, ena_flux_ub;
} orc_t;

static void *xc(size_t n) { void *p = calloc(n ? n : 1, 1); return p; }

/* ---- cell scans exactly as written in the reference ------------------------ */

/* _8_1_1_KineticUpdateA.f90:221-273 */
static int scan_qcell(const orc_cfg *c, double q, int prev) {
  int nq = c->nq, out = prev;
  if (c->qGL[0] <= 0) {
    for (int Q = 0; Q < nq; ++Q) {
      if (Q == 0 && c->qGL[Q] <= q && q <= c->qGH[Q]) { out = Q + 1; break; }
      if (Q != 0 && c->qGL[Q] < q && q <= c->qGH[Q]) { out = Q + 1; break; }
    }
    if (c->qGL[0] > q) out = 0;
    else if (c->qGH[nq - 1] < q) out = -1;
  } else {
    for (int Q = 0; Q < nq; ++Q) {
      if (Q == 0 && c->qGL[Q] >= q && q >= c->qGH[Q]) { out = Q + 1; break; }
      if (Q != 0 && c->qGL[Q] > q && q >= c->qGH[Q]) { out = Q + 1; break; }
    }
    if (c->qGL[0] < q) out = 0;
    else if (c->qGH[nq - 1] > q) out = -1;
  }
  if (isnan(q)) out = 0;    /* documented: NaN positions leave through the LB */
  return out;
}

/* 2-D velocity scan, _8_1_1_KineticUpdateA.f90:1087-1127: no match keeps prev. */
static int scan_vcell(const double *GL, const double *GH, int n, double v, int prev) {
  for (int i = 0; i < n; ++i) {
    if (i == 0 && GL[i] <= v && v <= GH[i]) return 1;
    if (i != 0 && GL[i] < v && v <= GH[i]) return i + 1;
  }
  return prev;
}

/* ---- fields ---------------------------------------------------------------- */

/* Fields of a master step from the previous master step's density (shared
 * per-tube arithmetic kaos_ambipolar/kaos_grav_cell). */
static void ambipolar_field(const orc_cfg *c, const double *M0prev, double *M0F, double *EA) {
  kaos_ambipolar(c->nq, c->qGL[0] <= 0, c->eamb, c->eambsign, c->filter_pt, c->qs,
                 c->hqC, c->dqC, c->Te, M0prev, M0F, EA);
}
/* M1ParFiltAvrgRT(nn): the moving-average filter of the previous master
 * step's M1Par (_8_7_AmbipolarEfield.f90:64-80), window M1ParMAfilterPt. */
static void m1par_filter(const orc_cfg *c, const double *Mprev, double *M1F) {
  kaos_moment_filter(Mprev + 2 * c->nq, M1F, c->nq, c->filter_pt_m1par ? c->filter_pt_m1par : c->filter_pt);
}
static void grav_field(const orc_cfg *c, double *EG) {
  for (int Q = 0; Q < c->nq; ++Q)
    EG[Q] = kaos_grav_cell(c->grav, c->ms, c->qs, c->cthGC[Q], c->rGC[Q], c->ellGC[Q]);
}

/* Per-particle interpolation at qq (_8_7_AmbipolarEfield.f90:1453-2720,
 * _8_7_2_Gravfield.f90:132-...): Qind-outer, particle-inner scan over the
 * half cells and centre intervals.  No match leaves *out unchanged.        */
static void interp_scan(const orc_cfg *c, const double *E, double qq, double *out, int kind) {
  int nq = c->nq, ub = nq - 1;
  double v = 0.0; int hit = 0;
  for (int Q = 0; Q < nq && !hit; ++Q) {
    if (c->qGL[0] <= 0) {
      if (Q == 0 && c->qGL[Q] <= qq && qq <= c->qGC[Q]) { v = E[Q]; hit = 1; }
      if (Q != 0 && Q != ub && c->qGC[Q - 1] < qq && qq <= c->qGC[Q]) {
        v = E[Q - 1] + (((qq - c->qGC[Q - 1]) / (c->qGC[Q] - c->qGC[Q - 1])) * (E[Q] - E[Q - 1])); hit = 1; }
      if (Q == ub) {
        if (c->qGC[Q - 1] < qq && qq < c->qGC[Q]) {
          v = E[Q - 1] + (((qq - c->qGC[Q - 1]) / (c->qGC[Q] - c->qGC[Q - 1])) * (E[Q] - E[Q - 1])); hit = 1; }
        if (c->qGC[Q] <= qq && qq <= c->qGH[Q]) { v = E[Q]; hit = 1; }
      }
    } else {
      if (Q == 0 && c->qGL[Q] >= qq && qq >= c->qGC[Q]) { v = E[Q]; hit = 1; }
      if (Q != 0 && Q != ub && c->qGC[Q - 1] > qq && qq >= c->qGC[Q]) {
        v = E[Q - 1] + (((qq - c->qGC[Q - 1]) / (c->qGC[Q] - c->qGC[Q - 1])) * (E[Q] - E[Q - 1])); hit = 1; }
      if (Q == ub) {
        if (c->qGC[Q - 1] > qq && qq > c->qGC[Q]) {
          v = E[Q - 1] + (((qq - c->qGC[Q - 1]) / (c->qGC[Q] - c->qGC[Q - 1])) * (E[Q] - E[Q - 1])); hit = 1; }
        if (c->qGC[Q] >= qq && qq >= c->qGH[Q]) { v = E[Q]; hit = 1; }
      }
    }
  }
  if (!hit) return;
  double qm = c->qs / c->ms;
  if (kind == 0) *out = kaos_aea_of(v, qm, c->eambsign);   /* AEAmagN */
  else *out = kaos_ag_of(v, qm);                           /* AGmagN */
}

/* ---- counts and moments ------------------------------------------------------ */

/* IonParticleCountsSub 2-D branch (_8_3_1_IonParticleCounts.f90:33-137,
 * 566-730): counts per config cell and per (config, Vperp, Vpar) cell of the
 * ions with a valid Qindk1, the phase-space count requiring both velocity
 * indices; then DistributionFunctions + moments per cell.                  */
static void counts_and_moments(orc_t *o, const int *qi, const int *vpi, const int *vri, int64_t ns) {
  orc_cfg *c = &o->c;
  int nq = c->nq, np = c->nvpar, nr = c->nvperp;
  int64_t ncell = (int64_t)nr * np;
  memset(o->nq_u, 0, sizeof(uint32_t) * nq);
  memset(o->nph_u, 0, sizeof(uint32_t) * nq * ncell);
  for (int64_t j = 0; j < ns; ++j) {
    int Q = qi[j];
    if (Q < 1 || Q > nq) continue;
    o->nq_u[Q - 1] += 1;
    int a = vri[j], b = vpi[j];
    if (a >= 1 && a <= nr && b >= 1 && b <= np)
      o->nph_u[(int64_t)(Q - 1) * ncell + (int64_t)(a - 1) * np + (b - 1)] += 1;
  }
  for (int Q = 0; Q < nq; ++Q) o->nq_counts[Q] = (double)o->nq_u[Q];
  for (int64_t i = 0; i < nq * ncell; ++i) o->nph_counts[i] = (double)o->nph_u[i];
  kaos_vgrid_t v = { np, nr, c->d3v, c->dvperp, c->dvpar, c->vperpGC, c->vparGC };
  const int G = kaos_mom_groups(np);
  double *A = (double *)xc(sizeof(double) * 3 * np), *x = (double *)xc(sizeof(double) * 6 * np);
  for (int Q = 0; Q < nq; ++Q) {
    double k = c->nsnormfac / c->d3xC[Q];
    o->NqRT[Q] = o->nq_counts[Q] * k;
    const uint32_t *N = o->nph_u + (int64_t)Q * ncell;
    /* column aggregates: row groups g = 0..G-1 added in order */
    for (int b = 0; b < np; ++b) {
      double P[3];
      uint32_t cs;
      kaos_mom_part(&v, N, b, 0, G, 0, v.nvperp - 1, k, &A[3 * b], &cs);
      for (int g = 1; g < G; ++g) {
        kaos_mom_part(&v, N, b, g, G, 0, v.nvperp - 1, k, P, &cs);
        for (int t = 0; t < 3; ++t) A[3 * b + t] = A[3 * b + t] + P[t];
      }
      double o4[4];
      kaos_mom_col_first(c->vparGC[b], &A[3 * b], o4);
      for (int t = 0; t < 4; ++t) x[t * np + b] = o4[t];
    }
    double S[4], Mq[6];
    for (int t = 0; t < 4; ++t) { kaos_tree_sum(&x[t * np], np); S[t] = x[t * np]; }
    kaos_mom_finish1(S, c->ms, Mq);
    for (int b = 0; b < np; ++b) {
      double o2[2];
      kaos_mom_col_second(c->vparGC[b], &A[3 * b], Mq[2], o2);
      x[4 * np + b] = o2[0]; x[5 * np + b] = o2[1];
    }
    double S2[2];
    for (int t = 0; t < 2; ++t) { kaos_tree_sum(&x[(4 + t) * np], np); S2[t] = x[(4 + t) * np]; }
    kaos_mom_finish2(S2, c->ms, Mq);
    for (int t = 0; t < 6; ++t) o->M[t * nq + Q] = Mq[t];
  }
  free(A);
  free(x);
}

/* 3-D velocity scan, _8_1_1_KineticUpdateA.f90:985-1080: 0 outside the grid
 * (no stale index), else the first cell: cell 1 [GL,GH], others (GL,GH]. */
static int scan_vcell3(const double *GL, const double *GH, int n, double v) {
  if (GL[0] > v || v > GH[n - 1]) return 0;
  for (int i = 0; i < n; ++i) {
    if (i == 0 && GL[i] <= v && v <= GH[i]) return 1;
    if (i != 0 && GL[i] < v && v <= GH[i]) return i + 1;
  }
  return 0;
}

/* IonParticleCountsSub (3-D branch, _8_3_1:300-361) and the moments of
 * _8_6_1.._8_6_6 (3-D branches) in the fixed order of kaos_mom3_part. */
static void counts3_and_moments(orc_t *o, int64_t ns) {
  orc_cfg *c = &o->c;
  int nq = c->nq, np = c->nvpar, n1 = c->nv1, n2 = c->nv2;
  int64_t ncell = (int64_t)n1 * n2 * np;
  memset(o->nq_u, 0, sizeof(uint32_t) * nq);
  memset(o->n3_u, 0, sizeof(uint32_t) * nq * ncell);
  for (int64_t j = 0; j < ns; ++j) {
    int Q = o->qind[j];
    if (Q < 1 || Q > nq) continue;
    o->nq_u[Q - 1] += 1;
    int a1 = o->i1ind[j], a2 = o->i2ind[j], b = o->ip3ind[j];
    if (a1 >= 1 && a2 >= 1 && b >= 1)
      o->n3_u[(int64_t)(Q - 1) * ncell + ((int64_t)(a1 - 1) * n2 + (a2 - 1)) * np + (b - 1)] += 1;
  }
  for (int Q = 0; Q < nq; ++Q) o->nq_counts[Q] = (double)o->nq_u[Q];
  kaos_vgrid3_t v = { np, n1, n2, c->d3v3, c->dv1, c->dv2, c->dvpar, c->v1GC, c->v2GC, c->vparGC };
  const int G = kaos_mom3_groups(np);
  double *A = (double *)xc(sizeof(double) * 5 * np), *x = (double *)xc(sizeof(double) * 8 * np);
  for (int Q = 0; Q < nq; ++Q) {
    double k = c->nsnormfac / c->d3xC[Q];
    o->NqRT[Q] = o->nq_counts[Q] * k;
    const uint32_t *N = o->n3_u + (int64_t)Q * ncell;
    for (int b = 0; b < np; ++b) {
      double P[5];
      uint32_t cs;
      kaos_mom3_part(&v, N, b, 0, G, 0, n1 * n2 - 1, k, &A[5 * b], &cs);
      for (int g = 1; g < G; ++g) {
        kaos_mom3_part(&v, N, b, g, G, 0, n1 * n2 - 1, k, P, &cs);
        for (int t = 0; t < 5; ++t) A[5 * b + t] = A[5 * b + t] + P[t];
      }
      double o6[6];
      kaos_mom3_col_first(c->vparGC[b], &A[5 * b], o6);
      for (int t = 0; t < 6; ++t) x[t * np + b] = o6[t];
    }
    double S[6], M8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < 6; ++t) { kaos_tree_sum(&x[t * np], np); S[t] = x[t * np]; }
    kaos_mom3_finish1(S, c->ms, M8);
    for (int b = 0; b < np; ++b) {
      double o2[2];
      kaos_mom3_col_second(c->vparGC[b], &A[5 * b], M8[3], o2);
      x[6 * np + b] = o2[0]; x[7 * np + b] = o2[1];
    }
    double S2[2];
    for (int t = 0; t < 2; ++t) { kaos_tree_sum(&x[(6 + t) * np], np); S2[t] = x[(6 + t) * np]; }
    kaos_mom3_finish2(S2, c->ms, M8);
    for (int t = 0; t < 8; ++t) o->M3[t * nq + Q] = M8[t];
    /* the 2-D moment block of this master step: the quantities both branches
     * define (M0, M1Par, M2, M2Par; M1Perp and M2Perp are 2-D only) */
    o->M[0 * nq + Q] = M8[0]; o->M[1 * nq + Q] = 0.0; o->M[2 * nq + Q] = M8[3];
    o->M[3 * nq + Q] = M8[4]; o->M[4 * nq + Q] = 0.0; o->M[5 * nq + Q] = M8[7];
  }
  free(A);
  free(x);
}

/* The 3-D velocity cells of ion j at velocity (vpar, v1, v2). */
static void vcells3(orc_t *o, int64_t j, double vpar, double v1, double v2) {
  orc_cfg *c = &o->c;
  o->ip3ind[j] = scan_vcell3(c->vparGL, c->vparGH, c->nvpar, vpar);
  o->i1ind[j] = scan_vcell3(c->v1GL, c->v1GH, c->nv1, v1);
  o->i2ind[j] = scan_vcell3(c->v2GL, c->v2GH, c->nv2, v2);
}

/* ---- state management ------------------------------------------------------ */

#define ARR_D(X) X(rs) X(vs) X(vperp) X(aea) X(ag) X(aeaN) X(agN) X(qk4) X(q0) X(vpar0) X(v1) X(v2) X(ck4)
#define ARR_I(X) X(qind) X(vparind) X(vperpind) X(i1ind) X(i2ind) X(ip3ind)
#define ARR_L(X) X(id)

void *orc_create(const orc_cfg *cfg, int64_t ns, const double *q0, const double *vpar0,
                 const double *vperp0, int64_t cap) {
  orc_t *o = (orc_t *)xc(sizeof(orc_t));
  o->c = *cfg;
  o->ns = ns; o->cap = cap < ns ? ns : cap;
#define AL(n) o->n = (double *)xc(sizeof(double) * o->cap);
  ARR_D(AL)
#undef AL
#define AL(n) o->n = (int *)xc(sizeof(int) * o->cap);
  ARR_I(AL)
#undef AL
#define AL(n) o->n = (int64_t *)xc(sizeof(int64_t) * o->cap);
  ARR_L(AL)
#undef AL
  int nq = cfg->nq;
  int64_t ncell = (int64_t)cfg->nvperp * cfg->nvpar;
  o->nq_counts = (double *)xc(sizeof(double) * nq);
  o->nph_counts = (double *)xc(sizeof(double) * nq * ncell);
  o->nq_u = (uint32_t *)xc(sizeof(uint32_t) * nq);
  o->nph_u = (uint32_t *)xc(sizeof(uint32_t) * nq * ncell);
  o->NqRT = (double *)xc(sizeof(double) * nq);
  o->M = (double *)xc(sizeof(double) * 6 * nq);
  o->Mprev = (double *)xc(sizeof(double) * 6 * nq);
  o->EA = (double *)xc(sizeof(double) * nq);
  o->EG = (double *)xc(sizeof(double) * nq);
  o->M0F = (double *)xc(sizeof(double) * nq);
  o->M1F = (double *)xc(sizeof(double) * nq);
  if (cfg->v3d) {
    o->n3_u = (uint32_t *)xc(sizeof(uint32_t) * nq * (int64_t)cfg->nv1 * cfg->nv2 * cfg->nvpar);
    o->M3 = (double *)xc(sizeof(double) * 8 * nq);
  }
  if (cfg->cx) {
    int64_t ec = cfg->ena_cap > 0 ? cfg->ena_cap : 1;
    o->cx_sigma = (double *)xc(sizeof(double) * nq);
    o->cx_nu = (double *)xc(sizeof(double) * nq);
    o->cx_poi = (double *)xc(sizeof(double) * nq);
    o->nq_ena = (uint32_t *)xc(sizeof(uint32_t) * nq);
    o->ex = (double *)xc(sizeof(double) * ec); o->ey = (double *)xc(sizeof(double) * ec);
    o->ez = (double *)xc(sizeof(double) * ec); o->evx = (double *)xc(sizeof(double) * ec);
    o->evy = (double *)xc(sizeof(double) * ec); o->evz = (double *)xc(sizeof(double) * ec);
    o->eid = (int64_t *)xc(sizeof(int64_t) * ec);
    o->eq = (int *)xc(sizeof(int) * ec);
    o->evp = (int *)xc(sizeof(int) * ec); o->evq = (int *)xc(sizeof(int) * ec); o->evphi = (int *)xc(sizeof(int) * ec);
    o->ena_hist = (uint32_t *)xc(sizeof(uint32_t) * nq * (int64_t)cfg->nvp * cfg->nvq * cfg->nvphi);
    o->ena_M = (double *)xc(sizeof(double) * 8 * nq);
  }
  /* initial conditions: x0 = root(q0), V0 = Vpar0 b-hat(x0) */
  for (int64_t j = 0; j < ns; ++j) {
    o->rs[j] = kaos_rs_of_q(q0[j], cfg->p);
    o->vs[j] = vpar0[j];
    o->vperp[j] = vperp0[j];
    o->v1[j] = vperp0[j];        /* gyrophase 0 unless orc_set_vperp sets both components */
    o->v2[j] = 0.0;
    o->q0[j] = q0[j];
    o->vpar0[j] = vpar0[j];
    o->id[j] = j;
  }
  o->next_id = ns;
  return o;
}

void orc_destroy(void *h) {
  orc_t *o = (orc_t *)h;
  if (!o) return;
#define FR(n) free(o->n);
  ARR_D(FR) ARR_I(FR) ARR_L(FR)
#undef FR
  free(o->nq_counts); free(o->nph_counts); free(o->nq_u); free(o->nph_u); free(o->NqRT); free(o->M); free(o->Mprev);
  free(o->EA); free(o->EG); free(o->M0F); free(o->M1F); free(o->n3_u); free(o->M3); free(o->EAin);
  free(o->cx_sigma); free(o->cx_nu); free(o->cx_poi); free(o->nq_ena);
  free(o->ex); free(o->ey); free(o->ez); free(o->evx); free(o->evy); free(o->evz); free(o->eid); free(o->eq);
  free(o->evp); free(o->evq); free(o->evphi); free(o->ena_hist); free(o->ena_M);
  free(o);
}

/* BoundaryConditionsSub at a master step: escape counts and fluxes
 * (_8_0_7_BoundaryConditions.f90:158-285) and stable pack (BCReIndexSub). */
static void bc_remove(orc_t *o) {
  orc_cfg *c = &o->c;
  int64_t k = 0, nlb = 0, nub = 0;
  for (int64_t j = 0; j < o->ns; ++j) {
    if (o->qind[j] == 0) { ++nlb; continue; }
    if (o->qind[j] == -1) { ++nub; continue; }
    if (k != j) {
#define MV(n) o->n[k] = o->n[j];
      ARR_D(MV) ARR_I(MV) ARR_L(MV)
#undef MV
    }
    ++k;
  }
  o->ns = k;
  o->esc_lb = (double)nlb; o->esc_ub = (double)nub;
  int ub = c->nq - 1;
  double dtm = (double)c->ndatfac * c->h;
  double NLB = (double)nlb * c->nsnormfac, NUB = (double)nub * c->nsnormfac;
  o->flux_lb = NLB / ((c->d3xC[0] / (c->hqC[0] * c->dqC[0])) * dtm);
  o->flux_ub = NUB / ((c->d3xC[ub] / (c->hqC[ub] * c->dqC[ub])) * dtm);
}

/* LB/UB ion injection after the master push (KSB:195-793). */
static int inject(orc_t *o, int64_t n, const double *uni) {
  orc_cfg *c = &o->c;
  int64_t nl = c->lbcond ? c->n_inj_lb : 0, nu = c->ubcond ? c->n_inj_ub : 0;
  if (o->ns + nl + nu > o->cap) return 2;
  for (int64_t k = 0; k < nl + nu; ++k) {
    int lb = (k < nl);
    int Qb = lb ? 0 : c->nq - 1;
    int64_t j = o->ns + k;
    double qc = c->qGC[Qb];
    double u[6];
    if (uni) memcpy(u, uni + 6 * k, sizeof(u));
    else kaos_inject_uniforms(c->seed, n, c->tube, k, u);
    o->qind[j] = scan_qcell(c, qc, 0);
    o->rs[j] = kaos_rs_of_q(qc, c->p);
    kaos_line_t L = kaos_line(o->rs[j], c->p);
    double Ts = lb ? c->Ts_lb : c->Ts_ub;
    double sigma = sqrt((KAOS_KB * Ts) / c->ms);
    double vpar, v1, v2, vperp;
    kaos_mb_project(u, sigma, c->cphi, c->sphi, L, &vpar, &v1, &v2, &vperp);
    o->vs[j] = vpar;
    o->vperp[j] = vperp;
    o->v1[j] = v1;
    o->v2[j] = v2;
    o->id[j] = o->next_id + k;
    o->vparind[j] = scan_vcell(c->vparGL, c->vparGH, c->nvpar, vpar, 0);
    o->vperpind[j] = scan_vcell(c->vperpGL, c->vperpGH, c->nvperp, vperp, 0);
    if (c->v3d) vcells3(o, j, vpar, v1, v2);   /* indices at injection, KSB:339-476 */
    /* A new slot's fields and stage-4 q are never written before its first
     * push: DataTypeReAllocA copies only the old ions (_8_0_6_DataTypeReAllocA.
     * f90:259-313) and the injection sets neither.  The field scans at the next
     * master step match no cell at an unset qk4 = 0 (_8_7_2_Gravfield.f90:305-
     * 420, AmbipolarEfield likewise) and leave AGmagN/AEAmagN unset, so the
     * first push after injection reads uninitialised memory for both fields.
     * The compiled reference reads 0 there for most batches (stale values of
     * other slots for some); 0 is taken here, with which the reference's own
     * loop is reproduced over four master steps. */
    o->aea[j] = o->aeaN[j] = 0.0;
    o->ag[j] = o->agN[j] = 0.0;
    o->qk4[j] = 0.0;
    o->ck4[j] = L.c;     /* a conversion before the first push: the frame of the injection point */
  }
  o->ns += nl + nu;
  o->next_id += nl + nu;
  o->injected = nl + nu;
  return 0;
}

/* Heating parameters of the tube; the heating region's arc length is the sum
 * of dsICRGT = hqC dqC over the cells (_1_SimParameterization.f90:1185-1186,
 * summed in WaveHeatingSub, _8_1_1_0_WaveHeating.f90:100). */
static kaos_icr_t icr_params(const orc_cfg *c) {
  kaos_icr_t P;
  P.qs = c->qs; P.ms = c->ms; P.h = c->h;
  P.lambda_perp = c->icr_lambda_perp; P.eta_lh = c->icr_eta_lh;
  P.xi1 = c->icr_xi_perp1; P.xi2 = c->icr_xi_perp2; P.s0 = c->icr_s0;
  P.omega_g0 = KAOS_TWOPI * c->icr_f0;
  P.chi1 = c->icr_chi_perp1; P.chi2 = c->icr_chi_perp2;
  double s = 0.0;
  for (int i = 0; i < c->nq; ++i) s = s + c->hqC[i] * c->dqC[i];
  P.sum_ds = s;
  return P;
}

/* KineticRK4UpdateSub over the live ions; with ICR on, WaveHeatingSub's kick
 * at stage A for ions inside the grid (_8_1_1_KineticUpdateA.f90:1392,
 * _8_1_1_0_WaveHeating.f90:32-33), after the pre-push Vperp cell is taken. */
static void push_all(orc_t *o, int64_t n) {
  orc_cfg *c = &o->c;
  kaos_push_par_t P = kaos_push_par(c->h, c->p, c->cphi, c->sphi, c->grav, c->eamb, c->eambsign);
  kaos_icr_t IP = icr_params(c);
  for (int64_t j = 0; j < o->ns; ++j) {
    kaos_line_t L = kaos_line(o->rs[j], c->p);
    kaos_qb_t gA = kaos_stageA(L);
    int qs = scan_qcell(c, gA.q, o->qind[j]);    /* the reference's linear scan */
    kaos_push_out_t r;
    kaos_push_ion(L, gA, qs, o->vs[j], o->ag[j], o->aea[j], &P, &r);
    o->ck4[j] = r.zC / sqrt(r.u2C);               /* cos(thetak4), KUC */
    o->qind[j] = qs;
    o->vparind[j] = scan_vcell(c->vparGL, c->vparGH, c->nvpar, r.vpar, o->vparind[j]);
    o->vperpind[j] = scan_vcell(c->vperpGL, c->vperpGH, c->nvperp, o->vperp[j], o->vperpind[j]);
    if (c->v3d) vcells3(o, j, r.vpar, o->v1[j], o->v2[j]);   /* pre-kick components */
    if (c->icr && qs != 0 && qs != -1) {
      double u[4], a1, a2, ap;
      kaos_icr_uniforms(c->seed, n, c->tube, o->id[j], u);
      kaos_icr_kick(&IP, L.r, L.w, r.vpar, o->v1[j], o->v2[j], u, &a1, &a2, &ap);
      o->v1[j] = a1; o->v2[j] = a2; o->vperp[j] = ap;
    }
    o->rs[j] = r.rsN;
    o->vs[j] = r.vparN;
    o->qk4[j] = r.qk4;
  }
}

/* ---- charge exchange -------------------------------------------------- */

/* Collision counts of every cell at a master step n (IonNeutralCollision-
 * FrequencyASub + the one-rank share of ...BSub, kaos_cx_cell), from the 3-D
 * moments of this master step, filtered with MOMENTFILTERflagT = 1
 * (_8_6_6_1:22-86: M1Par with its own window, M1Perp1, M1Perp2, M2Par with
 * theirs). */
static void cx_freq(orc_t *o, int64_t n) {
  orc_cfg *c = &o->c;
  int nq = c->nq;
  const double *M3 = o->M3;
  double *f = (double *)xc(sizeof(double) * 4 * nq);
  const double *m1par = M3 + 3 * nq, *m1p1 = M3 + 1 * nq, *m1p2 = M3 + 2 * nq, *m2par = M3 + 7 * nq;
  if (c->cx_filter) {
    int w1 = c->filter_pt_m1par ? c->filter_pt_m1par : c->filter_pt;
    kaos_moment_filter(m1par, f, nq, w1);
    kaos_moment_filter(m1p1, f + nq, nq, c->cx_filter_pt);
    kaos_moment_filter(m1p2, f + 2 * nq, nq, c->cx_filter_pt);
    kaos_moment_filter(m2par, f + 3 * nq, nq, c->cx_filter_pt);
    m1par = f; m1p1 = f + nq; m1p2 = f + 2 * nq; m2par = f + 3 * nq;
  }
  const double sig_n = sqrt((KAOS_KB * c->t_neut) / c->m_neut);
  for (int Q = 0; Q < nq; ++Q) {
    kaos_cx_in_t I;
    I.m0 = M3[Q]; I.m1par = m1par[Q]; I.m1p1 = m1p1[Q]; I.m1p2 = m1p2[Q]; I.m2par = m2par[Q];
    if (c->cx_filter && I.m0 == 0.0) I.m2par = 0.0;   /* sigma is set only for populated cells */
    I.cth = c->cthGC[Q]; I.ell = c->ellGC[Q]; I.cphi = c->cphi; I.sphi = c->sphi;
    I.nsneut = c->nsneut[Q]; I.h = c->h; I.sig_n = sig_n;
    I.NqRT = o->NqRT[Q]; I.d3x = c->d3xC[Q]; I.nsnormfac = c->nsnormfac; I.nqc = o->nq_counts[Q];
    I.sig_scale = c->cx_sigma_scale;
    kaos_cx_out_t r = kaos_cx_cell(&I, c->seed, n, c->tube, Q + 1);
    o->cx_sigma[Q] = r.sigma; o->cx_nu[Q] = r.nu; o->cx_poi[Q] = r.poi;
  }
  free(f);
}

/* ENA production at a step n >= 2 (_8_KineticSolver.f90:95-166): in cell
 * order, the first nint(nuIonNeutPoiT) ions (array order = id order) whose
 * Qindk1 is the cell become ENAs (IonNeutralCollisionsSub); they leave the
 * ion arrays (stable pack) for the ENA arrays. */
static int cx_convert(orc_t *o) {
  orc_cfg *c = &o->c;
  int nq = c->nq;
  int64_t *need = (int64_t *)xc(sizeof(int64_t) * nq);
  int64_t tot = 0;
  for (int Q = 0; Q < nq; ++Q) { need[Q] = kaos_cx_take(o->cx_poi[Q]); tot += need[Q]; }
  if (!tot) { free(need); return 0; }
  char *conv = (char *)xc((size_t)o->ns);
  for (int64_t j = 0; j < o->ns; ++j) {
    int Q = o->qind[j];
    if (Q >= 1 && need[Q - 1] > 0) { conv[j] = 1; --need[Q - 1]; }
  }
  int64_t k = 0;
  for (int64_t j = 0; j < o->ns; ++j) {
    if (conv[j]) {
      if (o->ens >= c->ena_cap) { free(conv); free(need); return 3; }
      double X[6];
      kaos_ena_from_ion(o->rs[j], o->vs[j], o->v1[j], o->v2[j], o->ck4[j], c->p, c->cphi, c->sphi, X);
      int64_t e = o->ens++;
      o->ex[e] = X[0]; o->ey[e] = X[1]; o->ez[e] = X[2];
      o->evx[e] = X[3]; o->evy[e] = X[4]; o->evz[e] = X[5];
      o->eid[e] = o->id[j];
      o->eq[e] = o->qind[j];
      o->evp[e] = o->evq[e] = o->evphi[e] = 0;
      ++o->ena_made;
      continue;
    }
    if (k != j) {
#define MV(n) o->n[k] = o->n[j];
      ARR_D(MV) ARR_I(MV) ARR_L(MV)
#undef MV
    }
    ++k;
  }
  o->ns = k;
  free(conv); free(need);
  return 0;
}

/* ENA escapes at a master step (_8_0_7_BoundaryConditions.f90:309-470):
 * ENAs whose Qindk1 is 0 (LB) or -1 (UB) are counted and removed. */
static void ena_remove(orc_t *o) {
  orc_cfg *c = &o->c;
  int64_t k = 0, nlb = 0, nub = 0;
  for (int64_t e = 0; e < o->ens; ++e) {
    if (o->eq[e] == 0) { ++nlb; continue; }
    if (o->eq[e] == -1) { ++nub; continue; }
    o->ex[k] = o->ex[e]; o->ey[k] = o->ey[e]; o->ez[k] = o->ez[e];
    o->evx[k] = o->evx[e]; o->evy[k] = o->evy[e]; o->evz[k] = o->evz[e];
    o->eid[k] = o->eid[e]; o->eq[k] = o->eq[e];
    o->evp[k] = o->evp[e]; o->evq[k] = o->evq[e]; o->evphi[k] = o->evphi[e];
    ++k;
  }
  o->ens = k;
  int ub = c->nq - 1;
  double dtm = (double)c->ndatfac * c->h;
  o->ena_esc_lb = (double)nlb; o->ena_esc_ub = (double)nub;
  o->ena_flux_lb = ((double)nlb * c->nsnormfac) / ((c->d3xC[0] / (c->hqC[0] * c->dqC[0])) * dtm);
  o->ena_flux_ub = ((double)nub * c->nsnormfac) / ((c->d3xC[ub] / (c->hqC[ub] * c->dqC[ub])) * dtm);
}

/* ENADistributionFunctionsSub and the eight ENA moments of every cell from
 * its counts: partial sums of lanes 0..255 (kaos_ena_mom_part), then the
 * pairwise tree; zeros for a cell without ENAs (NqENART = 0). */
static void ena_moments(const orc_cfg *c, const uint32_t *nq_ena, const uint32_t *hist, double *M8) {
  const int nb = c->nvp * c->nvq * c->nvphi;
  kaos_enagrid_t g = {c->nvp, c->nvq, c->nvphi, c->dvp, c->dvq, c->dvphi, c->vpGC, c->vqGC, c->vphiGC,
                      c->ena_hh, c->ena_d33v};
  double P[4][KAOS_ENA_LANES];
  for (int Q = 0; Q < c->nq; ++Q) {
    g.hh = c->ena_hh + (int64_t)Q * nb;             /* the cell's metric row */
    g.d33v = c->ena_d33v + (int64_t)Q * nb;
    double *M = M8 + 8 * Q;
    for (int i = 0; i < 8; ++i) M[i] = 0.0;
    if (nq_ena[Q] == 0) continue;
    const uint32_t *cnt = hist + (int64_t)Q * nb;
    const double k = c->nsnormfac / c->d3xC[Q];
    double S[4], m1[3] = {0.0, 0.0, 0.0};
    for (int stage = 1; stage <= 2; ++stage) {
      for (int l = 0; l < KAOS_ENA_LANES; ++l) {
        double o4[4];
        kaos_ena_mom_part(&g, cnt, l, k, stage, m1, o4);
        for (int i = 0; i < 4; ++i) P[i][l] = o4[i];
      }
      for (int i = 0; i < 4; ++i) { kaos_tree_sum(P[i], KAOS_ENA_LANES); S[i] = P[i][0]; }
      if (stage == 1) { kaos_ena_mom_finish1(S, M); m1[0] = M[1]; m1[1] = M[2]; m1[2] = M[3]; }
      else kaos_ena_mom_finish2(S, c->ms, M);
    }
  }
}

/* The eight ENA moments of given counts (nq * bins) and per-cell totals, as
 * [8][nq] (for the pins). */
void orc_ena_moments_from_counts(const orc_cfg *c, const double *counts, const double *nq_ena, double *M8) {
  const int64_t n = (int64_t)c->nq * c->nvp * c->nvq * c->nvphi;
  uint32_t *h = (uint32_t *)xc(sizeof(uint32_t) * n), *q = (uint32_t *)xc(sizeof(uint32_t) * c->nq);
  double *M = (double *)xc(sizeof(double) * 8 * c->nq);
  for (int64_t i = 0; i < n; ++i) h[i] = (uint32_t)counts[i];
  for (int Q = 0; Q < c->nq; ++Q) q[Q] = (uint32_t)nq_ena[Q];
  ena_moments(c, q, h, M);
  for (int Q = 0; Q < c->nq; ++Q)
    for (int i = 0; i < 8; ++i) M8[(int64_t)i * c->nq + Q] = M[8 * Q + i];
  free(h); free(q); free(M);
}

void orc_ena_vcomp(const double *X, double *o3) { kaos_ena_vcomp(X, &o3[0], &o3[1], &o3[2]); }

/* The ENA push of a step: Qindk1 at the pre-push position (the ion scan of
 * KUA:221-273 on the ENA's q), counted at master steps (ENAParticleCountsSub,
 * _8_3_2:33-60), then the ballistic RK4. */
static void ena_push(orc_t *o, int master) {
  orc_cfg *c = &o->c;
  const int nb = c->nvp * c->nvq * c->nvphi;
  if (master) {
    memset(o->nq_ena, 0, sizeof(uint32_t) * c->nq);
    memset(o->ena_hist, 0, sizeof(uint32_t) * c->nq * (int64_t)nb);
  }
  for (int64_t e = 0; e < o->ens; ++e) {
    double X[6] = {o->ex[e], o->ey[e], o->ez[e], o->evx[e], o->evy[e], o->evz[e]};
    double q = kaos_ena_q(X[0], X[1], X[2]);
    int Q = (q == q) ? scan_qcell(c, q, o->eq[e]) : 0;
    o->eq[e] = Q;
    if (Q >= 1) {
      /* the ENA velocity cells at the pre-push state, stale off the grid (KUA:1156-1290) */
      double vp, vq, vphi;
      kaos_ena_vcomp(X, &vp, &vq, &vphi);
      o->evp[e] = scan_vcell(c->vpGL, c->vpGH, c->nvp, vp, o->evp[e]);
      o->evq[e] = scan_vcell(c->vqGL, c->vqGH, c->nvq, vq, o->evq[e]);
      o->evphi[e] = scan_vcell(c->vphiGL, c->vphiGH, c->nvphi, vphi, o->evphi[e]);
    }
    if (master && Q >= 1) {
      o->nq_ena[Q - 1]++;
      if (o->evp[e] >= 1 && o->evq[e] >= 1 && o->evphi[e] >= 1)
        o->ena_hist[(int64_t)(Q - 1) * nb + ((int64_t)(o->evp[e] - 1) * c->nvq + (o->evq[e] - 1)) * c->nvphi +
                    (o->evphi[e] - 1)]++;
    }
    kaos_ena_push(X, c->h, c->grav);
    o->ex[e] = X[0]; o->ey[e] = X[1]; o->ez[e] = X[2];
    o->evx[e] = X[3]; o->evy[e] = X[4]; o->evz[e] = X[5];
  }
  if (master) ena_moments(c, o->nq_ena, o->ena_hist, o->ena_M);
}

/* ENA phase-space counts [Q][ip][iq][iphi] and the eight ENA moments per cell
 * (M0, M1P, M1Q, M1PHI, M2, M2P, M2Q, M2PHI) of the last master step. */
void orc_get_ena_moments(void *h, double *counts, double *M8) {
  orc_t *o = (orc_t *)h;
  orc_cfg *c = &o->c;
  if (!c->cx) return;
  int64_t n = (int64_t)c->nq * c->nvp * c->nvq * c->nvphi;
  if (counts) for (int64_t i = 0; i < n; ++i) counts[i] = (double)o->ena_hist[i];
  if (M8) {
    for (int Q = 0; Q < c->nq; ++Q)
      for (int i = 0; i < 8; ++i) M8[(int64_t)i * c->nq + Q] = o->ena_M[8 * Q + i];
  }
}

void orc_get_cx(void *h, double *sigma, double *nu, double *poi) {
  orc_t *o = (orc_t *)h;
  int nq = o->c.nq;
  if (!o->c.cx) return;
  if (sigma) memcpy(sigma, o->cx_sigma, sizeof(double) * nq);
  if (nu) memcpy(nu, o->cx_nu, sizeof(double) * nq);
  if (poi) memcpy(poi, o->cx_poi, sizeof(double) * nq);
}

int64_t orc_ena_count(void *h) { return ((orc_t *)h)->ens; }

/* The ENAs in id order: x, y, z, Vx, Vy, Vz, id, Qindk1. */
void orc_get_enas(void *h, double *X6, int64_t *id, int64_t *qind) {
  orc_t *o = (orc_t *)h;
  int64_t n = o->ens;
  int64_t *ord = (int64_t *)xc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t i = 0; i < n; ++i) ord[i] = i;
  /* insertion sort by id (few ENAs in tests) */
  for (int64_t i = 1; i < n; ++i) {
    int64_t v = ord[i], k = i - 1;
    while (k >= 0 && o->eid[ord[k]] > o->eid[v]) { ord[k + 1] = ord[k]; --k; }
    ord[k + 1] = v;
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t e = ord[i];
    double v[6] = {o->ex[e], o->ey[e], o->ez[e], o->evx[e], o->evy[e], o->evz[e]};
    for (int k = 0; k < 6; ++k) X6[6 * i + k] = v[k];
    id[i] = o->eid[e];
    qind[i] = o->eq[e];
  }
  free(ord);
}

/* ENA counts NqReNormENAT of the last master step and the LB/UB ENA escapes
 * (counts, then fluxes). */
void orc_get_ena_master(void *h, double *nq_ena, double *esc4) {
  orc_t *o = (orc_t *)h;
  if (!o->c.cx) return;
  for (int Q = 0; Q < o->c.nq; ++Q) nq_ena[Q] = (double)o->nq_ena[Q];
  esc4[0] = o->ena_esc_lb; esc4[1] = o->ena_esc_ub; esc4[2] = o->ena_flux_lb; esc4[3] = o->ena_flux_ub;
}

/* The input field of a restart replaces the computed EAmagRT after the
 * stencils and the LB override: at step 1 (modes 1, 2) and at every later
 * master step (mode 2) (_8_7_AmbipolarEfield.f90:1375-1407). */
static void field_input(orc_t *o, int first) {
  if (o->EAin && (o->ea_mode == 2 || (o->ea_mode == 1 && first)))
    memcpy(o->EA, o->EAin, sizeof(double) * o->c.nq);
}

void orc_set_field_input(void *h, const double *EA, int mode) {
  orc_t *o = (orc_t *)h;
  if (!o->EAin) o->EAin = (double *)xc(sizeof(double) * o->c.nq);
  if (EA) memcpy(o->EAin, EA, sizeof(double) * o->c.nq);
  o->ea_mode = mode;
}

/* One time step n (n = 1 is KineticSolverASub).  uniforms: 6 per injected
 * ion of this step (LB first), or NULL for the seeded Philox stream.        */
int orc_step(void *h, int64_t n, const double *uniforms) {
  orc_t *o = (orc_t *)h;
  orc_cfg *c = &o->c;
  int nq = c->nq;
  o->injected = 0;
  if (c->cx && n >= 2) {
    int rc = cx_convert(o);
    if (rc) return rc;
  }
  if (n == 1) {
    /* k0 indices of the initial state, with q0 clamped into the grid (KSA:55-320) */
    for (int64_t j = 0; j < o->ns; ++j) {
      double q = o->q0[j];
      if (c->qGL[0] <= 0) {
        if (q <= c->qGL[0]) q = c->qGL[0];
        else if (q > c->qGH[nq - 1]) q = c->qGH[nq - 1];
      } else {
        if (q > c->qGL[0]) q = c->qGL[0];
        else if (q < c->qGH[nq - 1]) q = c->qGH[nq - 1];
      }
      o->q0[j] = q;
      o->qind[j] = scan_qcell(c, q, 0);
      o->vparind[j] = scan_vcell(c->vparGL, c->vparGH, c->nvpar, o->vpar0[j], 0);
      o->vperpind[j] = scan_vcell(c->vperpGL, c->vperpGH, c->nvperp, o->vperp[j], 0);
      if (c->v3d) vcells3(o, j, o->vpar0[j], o->v1[j], o->v2[j]);   /* Vparindk0, Vperp1/2indk0 */
    }
    if (c->v3d) counts3_and_moments(o, o->ns);
    else counts_and_moments(o, o->qind, o->vparind, o->vperpind, o->ns);
    if (c->cx) cx_freq(o, 1);
    memset(o->Mprev, 0, sizeof(double) * 6 * nq);   /* moments of master step 0: none */
    ambipolar_field(c, o->Mprev, o->M0F, o->EA);
    field_input(o, 1);
    m1par_filter(c, o->Mprev, o->M1F);
    grav_field(c, o->EG);
    for (int64_t j = 0; j < o->ns; ++j) {
      interp_scan(c, o->EA, o->q0[j], &o->aeaN[j], 0);
      interp_scan(c, o->EG, o->q0[j], &o->agN[j], 1);
    }
    memcpy(o->Mprev, o->M, sizeof(double) * 6 * nq);
    o->nn = 1;
    o->master_last = 1;
  } else {
    int master = (n % c->ndatfac) == 0;
    if (master) {
      bc_remove(o);
      if (c->cx) ena_remove(o);
      ambipolar_field(c, o->Mprev, o->M0F, o->EA);
      field_input(o, 0);
      m1par_filter(c, o->Mprev, o->M1F);
      grav_field(c, o->EG);
      for (int64_t j = 0; j < o->ns; ++j) {
        interp_scan(c, o->EA, o->qk4[j], &o->aeaN[j], 0);
        interp_scan(c, o->EG, o->qk4[j], &o->agN[j], 1);
      }
    }
    o->master_last = master;
  }
  for (int64_t j = 0; j < o->ns; ++j) { o->aea[j] = o->aeaN[j]; o->ag[j] = o->agN[j]; }
  push_all(o, n);
  if (c->cx) ena_push(o, n == 1 || o->master_last);
  if (n > 1 && o->master_last) {
    int rc = inject(o, n, uniforms);
    if (rc) return rc;
    if (c->v3d) counts3_and_moments(o, o->ns);
    else counts_and_moments(o, o->qind, o->vparind, o->vperpind, o->ns);
    if (c->cx) cx_freq(o, n);
    memcpy(o->Mprev, o->M, sizeof(double) * 6 * nq);
    o->nn = n / c->ndatfac + 1;
  }
  return 0;
}

int64_t orc_count(void *h) { return ((orc_t *)h)->ns; }
int orc_master_last(void *h) { return ((orc_t *)h)->master_last; }

void orc_get_particles(void *h, double *x, double *y, double *z, double *Vx, double *Vy,
                       double *Vz, double *vperp, double *qk4, double *aea, double *ag,
                       int *qind, int *vparind, int *vperpind) {
  orc_t *o = (orc_t *)h;
  orc_cfg *c = &o->c;
  size_t nd = sizeof(double) * o->ns, ni = sizeof(int) * o->ns;
  for (int64_t j = 0; j < o->ns; ++j) {
    double a[6];
    kaos_cartesian(o->rs[j], o->vs[j], c->p, c->cphi, c->sphi, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]);
    if (x) x[j] = a[0];
    if (y) y[j] = a[1];
    if (z) z[j] = a[2];
    if (Vx) Vx[j] = a[3];
    if (Vy) Vy[j] = a[4];
    if (Vz) Vz[j] = a[5];
  }
  if (vperp) memcpy(vperp, o->vperp, nd);
  if (qk4) memcpy(qk4, o->qk4, nd);
  if (aea) memcpy(aea, o->aeaN, nd);
  if (ag) memcpy(ag, o->agN, nd);
  if (qind) memcpy(qind, o->qind, ni);
  if (vparind) memcpy(vparind, o->vparind, ni);
  if (vperpind) memcpy(vperpind, o->vperpind, ni);
}

/* Both perpendicular components of the uploaded ions (Vperp10T, Vperp20T;
 * KSA:389-391); Vperp = |(Vperp1, Vperp2)|. */
void orc_set_vperp(void *h, const double *v1, const double *v2) {
  orc_t *o = (orc_t *)h;
  for (int64_t j = 0; j < o->ns; ++j) {
    o->v1[j] = v1[j]; o->v2[j] = v2[j];
    o->vperp[j] = fabs(sqrt(v1[j] * v1[j] + v2[j] * v2[j]));
  }
}

/* The initial dipole coordinate q0 that step 1 clamps and scans for the k0
 * cells and interpolates the first fields at (KSA:55-320, Gravfield:132-300),
 * replaced after the positions were set from the q0 given to orc_create; a
 * test that feeds the reference's own step-1 inputs sets it. */
void orc_set_q0(void *h, const double *q0) {
  orc_t *o = (orc_t *)h;
  for (int64_t j = 0; j < o->ns; ++j) o->q0[j] = q0[j];
}

void orc_get_vperp12(void *h, double *v1, double *v2, int64_t *id) {
  orc_t *o = (orc_t *)h;
  if (v1) memcpy(v1, o->v1, sizeof(double) * o->ns);
  if (v2) memcpy(v2, o->v2, sizeof(double) * o->ns);
  if (id) memcpy(id, o->id, sizeof(int64_t) * o->ns);
}

/* Line state (signed radius, V||) of the live ions. */
void orc_get_lines(void *h, double *rs, double *vs) {
  orc_t *o = (orc_t *)h;
  memcpy(rs, o->rs, sizeof(double) * o->ns);
  memcpy(vs, o->vs, sizeof(double) * o->ns);
}

/* Outputs of the last master step. esc[0..3] = LB count, UB count, LB flux, UB flux. */
void orc_get_master(void *h, double *nq_counts, double *nph_counts, double *NqRT, double *M,
                    double *EA, double *EG, double *M0F, double *esc) {
  orc_t *o = (orc_t *)h;
  orc_cfg *c = &o->c;
  int nq = c->nq;
  int64_t ncell = (int64_t)c->nvperp * c->nvpar;
  if (nq_counts) memcpy(nq_counts, o->nq_counts, sizeof(double) * nq);
  if (nph_counts) memcpy(nph_counts, o->nph_counts, sizeof(double) * nq * ncell);
  if (NqRT) memcpy(NqRT, o->NqRT, sizeof(double) * nq);
  if (M) memcpy(M, o->M, sizeof(double) * 6 * nq);
  if (EA) memcpy(EA, o->EA, sizeof(double) * nq);
  if (EG) memcpy(EG, o->EG, sizeof(double) * nq);
  if (M0F) memcpy(M0F, o->M0F, sizeof(double) * nq);
  if (esc) { esc[0] = o->esc_lb; esc[1] = o->esc_ub; esc[2] = o->flux_lb; esc[3] = o->flux_ub; }
}

/* The 3-D outputs of the last master step: counts [Q][i1][i2][ipar] and the
 * 8 moments (M0, M1Perp1, M1Perp2, M1Par, M2, M2Perp1, M2Perp2, M2Par). */
void orc_get_master3(void *h, double *n3, double *M8) {
  orc_t *o = (orc_t *)h;
  orc_cfg *c = &o->c;
  int64_t n = (int64_t)c->nq * c->nv1 * c->nv2 * c->nvpar;
  if (n3) for (int64_t i = 0; i < n; ++i) n3[i] = (double)o->n3_u[i];
  if (M8) memcpy(M8, o->M3, sizeof(double) * 8 * c->nq);
}

void orc_get_m1f(void *h, double *M1F) {
  orc_t *o = (orc_t *)h;
  memcpy(M1F, o->M1F, sizeof(double) * o->c.nq);
}

/* ---- standalone pieces exposed for the unit tests --------------------------- */

void orc_filter(const double *in, double *out, int nq, int W) { kaos_moment_filter(in, out, nq, W); }
void orc_ambipolar(const orc_cfg *c, const double *M0prev, double *M0F, double *EA) {
  ambipolar_field(c, M0prev, M0F, EA);
}
void orc_gravity(const orc_cfg *c, double *EG) { grav_field(c, EG); }
void orc_ea_stencil(const orc_cfg *c, const double *M0prev, const double *nbar, double *EA) {
  kaos_ambipolar_stencil(c->nq, c->qGL[0] <= 0, c->eamb, c->eambsign, c->qs, c->hqC, c->dqC, c->Te,
                         M0prev, nbar, EA);
}
int orc_find_qcell(const orc_cfg *c, double q) {
  return kaos_find_qcell(q, c->qGH, c->qGL[0], c->qGH[c->nq - 1], c->nq, c->qGL[0] > 0);
}
int orc_find_vcell(const double *GL, const double *GH, int n, double v) { return kaos_find_vcell(v, GH, GL[0], n); }
int orc_interp_cells(const orc_cfg *c, const double *E, double q, double *out) {
  return kaos_interp_cells(q, c->qGC, E, c->nq, c->qGL[0], c->qGH[c->nq - 1], c->qGL[0] > 0, out);
}
double orc_interp_scan(const orc_cfg *c, const double *E, double q, double init, int kind) {
  double v = init; interp_scan(c, E, q, &v, kind); return v;
}
void orc_mb(const double *u, double sigma, double cphi, double sphi, double rs, double p, double *o4) {
  kaos_line_t L = kaos_line(rs, p);
  kaos_mb_project(u, sigma, cphi, sphi, L, &o4[0], &o4[1], &o4[2], &o4[3]);
}
double orc_gauss(double u1, double u2) { return kaos_gauss(u1, u2); }
int orc_scan_qcell(const orc_cfg *c, double q) { return scan_qcell(c, q, 0); }
int orc_scan_vcell(const double *GL, const double *GH, int n, double v, int prev) {
  return scan_vcell(GL, GH, n, v, prev);
}
/* Moments from given integer counts (for closed-form tests). */
void orc_moments_from_counts(const orc_cfg *cfg, const double *nq_counts, const double *nph_counts,
                             double *NqRT, double *M) {
  orc_t o;
  memset(&o, 0, sizeof(o));
  o.c = *cfg;
  int nq = cfg->nq;
  int64_t ncell = (int64_t)cfg->nvperp * cfg->nvpar;
  int64_t tot = 0;
  for (int Q = 0; Q < nq; ++Q) tot += (int64_t)nq_counts[Q];
  int *qi = (int *)xc(sizeof(int) * (tot + 1)), *vp = (int *)xc(sizeof(int) * (tot + 1)),
      *vr = (int *)xc(sizeof(int) * (tot + 1));
  int64_t k = 0;
  for (int Q = 0; Q < nq; ++Q) {
    int64_t inph = 0;
    for (int a = 0; a < cfg->nvperp; ++a)
      for (int b = 0; b < cfg->nvpar; ++b) {
        int64_t cnt = (int64_t)nph_counts[(int64_t)Q * ncell + a * cfg->nvpar + b];
        for (int64_t t = 0; t < cnt; ++t) { qi[k] = Q + 1; vr[k] = a + 1; vp[k] = b + 1; ++k; }
        inph += cnt;
      }
    for (int64_t t = inph; t < (int64_t)nq_counts[Q]; ++t) { qi[k] = Q + 1; vr[k] = 0; vp[k] = 0; ++k; }
  }
  o.nq_counts = (double *)xc(sizeof(double) * nq);
  o.nph_counts = (double *)xc(sizeof(double) * nq * ncell);
  o.nq_u = (uint32_t *)xc(sizeof(uint32_t) * nq);
  o.nph_u = (uint32_t *)xc(sizeof(uint32_t) * nq * ncell);
  o.NqRT = NqRT; o.M = M;
  counts_and_moments(&o, qi, vp, vr, k);
  free(qi); free(vp); free(vr); free(o.nq_counts); free(o.nph_counts); free(o.nq_u); free(o.nph_u);
}
/* One push of given line states (no cell bookkeeping), for the tests:
 * out[j] = (xN, yN, zN, VxN, VyN, VzN, rsN, VparN, Vpar, qk1, qk4). */
void orc_push_lines(const orc_cfg *c, int64_t n, const double *rs, const double *vs,
                    const double *ag, const double *aea, double *out) {
  kaos_push_par_t P = kaos_push_par(c->h, c->p, c->cphi, c->sphi, c->grav, c->eamb, c->eambsign);
  for (int64_t j = 0; j < n; ++j) {
    kaos_line_t L = kaos_line(rs[j], c->p);
    kaos_qb_t gA = kaos_stageA(L);
    int qs = scan_qcell(c, gA.q, 0);
    kaos_push_out_t r;
    kaos_push_ion(L, gA, qs, vs[j], ag[j], aea[j], &P, &r);
    double *w = out + 11 * j;
    kaos_cartesian(r.rsN, r.vparN, c->p, c->cphi, c->sphi, &w[0], &w[1], &w[2], &w[3], &w[4], &w[5]);
    w[6] = r.rsN; w[7] = r.vparN; w[8] = r.vpar; w[9] = r.qk1; w[10] = r.qk4;
  }
}
void orc_cartesian(const orc_cfg *c, int64_t n, const double *rs, const double *vs, double *out6) {
  for (int64_t j = 0; j < n; ++j)
    kaos_cartesian(rs[j], vs[j], c->p, c->cphi, c->sphi, &out6[6 * j], &out6[6 * j + 1], &out6[6 * j + 2],
                   &out6[6 * j + 3], &out6[6 * j + 4], &out6[6 * j + 5]);
}
int orc_find_qcell_from(const orc_cfg *c, double q, int hint) {
  return kaos_find_qcell_from(q, c->qGH, c->qGL[0], c->qGH[c->nq - 1], c->nq, c->qGL[0] > 0, hint);
}
int orc_find_vcell_guess(const double *GL, const double *GH, int n, double v) {
  return kaos_find_vcell_guess(v, GH, GL[0], n, (double)n / (GH[n - 1] - GL[0]));
}
int orc_interp2_from(const orc_cfg *c, const double *E, double q, int hint, double *out) {
  return kaos_interp2_from(q, c->qGC, E, (const double *)0, c->nq, c->qGL[0], c->qGH[c->nq - 1],
                           c->qGL[0] > 0, hint, out, (double *)0);
}
double orc_root_r(double q, double p) { return kaos_root_r(q, p); }
double orc_quartic_newton(double A, double B, double ip, double g0) { return kaos_quartic_newton(A, B, ip, g0); }
double orc_sinh_asinh_third(double y) { return kaos_sinh_asinh_third(y); }
double orc_log(double x) { return kaos_log(x); }
double orc_exp(double x) { return kaos_exp(x); }
double orc_sin2pi(double u) { return kaos_sin2pi(u); }
double orc_pow_neg(double x, double chi) { return kaos_pow_neg(x, chi); }
void orc_philox4x32(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { kaos_philox4x32(ctr, key, out); }
void orc_icr_uniforms(uint64_t seed, int64_t step, int tube, int64_t id, double *u) {
  kaos_icr_uniforms(seed, step, tube, id, u);
}
void orc_icr_kick(const orc_cfg *c, double r, double w, double vpar, double v1, double v2, const double *u,
                  double *out3) {
  kaos_icr_t P = icr_params(c);
  kaos_icr_kick(&P, r, w, vpar, v1, v2, u, &out3[0], &out3[1], &out3[2]);
}
double orc_cos2pi(double u) { return kaos_cos2pi(u); }
void orc_philox(const uint64_t *ctr, const uint64_t *key, uint64_t *out) { kaos_philox4x64(ctr, key, out); }
void orc_inject_uniforms(uint64_t seed, int64_t step, int tube, int64_t k, double *u) {
  kaos_inject_uniforms(seed, step, tube, k, u);
}

/* ---- initial ions ----------------------------------------------------------- */

/* NsFARRp of one cell: nsnormC0 = (n/nsnormfac) d3x truncated on its
 * assignment to the integer array (_KineticMainParams.f90:64,
 * _0_1_1_ConfigGridGeneratorB.f90:83-85, _2_1_DensityProfileA1.f90:45). */
int64_t orc_cell_count(double density, double d3x, double nsnormfac) {
  return kaos_cell_count(density, d3x, nsnormfac);
}

/* The initial ions of a tube as the reference builds them, loop for loop:
 * for each config cell (DensityProfileBSub, _4_DensityProfileB.f90:31-46),
 * counts[Q] ions at (pGC, qGC) of the cell; the root solve gives the signed
 * radius (DipolePolynomialSolverSub); each ion's MB velocity at the cell's
 * temperature (VelocityDistributionSub, _6_VelocityDistribution.f90:54-85),
 * its uniforms from the seeded stream of its running index i
 * (InitialConditionsSub's concatenation order, _7_InitialConditions.f90:72-87).
 * Returns the number of ions written. */
int64_t orc_seed_ions(const orc_cfg *c, const int64_t *counts, const double *temp, uint64_t seed, int tube,
                      double *q0, double *rs, double *vpar, double *v1, double *v2, double *vperp) {
  int64_t i = 0;
  for (int Q = 0; Q < c->nq; ++Q) {
    double sigma = kaos_thermal_speed(temp[Q], c->ms);
    for (int64_t k = 0; k < counts[Q]; ++k, ++i) {
      double u[6];
      kaos_seed_uniforms(seed, tube, i, u);
      q0[i] = c->qGC[Q];
      kaos_seed_ion(c->qGC[Q], c->p, sigma, c->cphi, c->sphi, u, &rs[i], &vpar[i], &v1[i], &v2[i], &vperp[i]);
    }
  }
  return i;
}
void orc_seed_uniforms(uint64_t seed, int tube, int64_t i, double *u) { kaos_seed_uniforms(seed, tube, i, u); }

/* Charge-exchange pieces for the pins (tests/test_oracle_cx.py). */
void orc_cx_block(uint64_t seed, int64_t step, int tube, int Q, uint64_t blk, double *u) {
  kaos_cx_block(seed, step, tube, Q, blk, u);
}
double orc_cx_sigma(double m2par) { return kaos_cx_sigma(m2par); }
double orc_poisson(double mu, uint64_t seed, int64_t step, int tube, int Q) {
  return kaos_poisson(mu, seed, step, tube, Q);
}
/* in: m0, m1par, m1p1, m1p2, m2par, cth, ell, cphi, sphi, nsneut, h, sig_n, NqRT, d3x, nsnormfac, nqc, sig_scale;
 * out: sigma, nu, poi */
void orc_cx_cell(const double *in, uint64_t seed, int64_t step, int tube, int Q, double *out) {
  kaos_cx_in_t I;
  I.m0 = in[0]; I.m1par = in[1]; I.m1p1 = in[2]; I.m1p2 = in[3]; I.m2par = in[4];
  I.cth = in[5]; I.ell = in[6]; I.cphi = in[7]; I.sphi = in[8]; I.nsneut = in[9]; I.h = in[10];
  I.sig_n = in[11]; I.NqRT = in[12]; I.d3x = in[13]; I.nsnormfac = in[14]; I.nqc = in[15];
  I.sig_scale = in[16];
  kaos_cx_out_t r = kaos_cx_cell(&I, seed, step, tube, Q);
  out[0] = r.sigma; out[1] = r.nu; out[2] = r.poi;
}
void orc_ena_from_ion(double rs, double vs, double v1, double v2, double ck4, double p, double cphi, double sphi,
                      double unused, double *X) {
  (void)unused;
  kaos_ena_from_ion(rs, vs, v1, v2, ck4, p, cphi, sphi, X);
}
void orc_ena_push(double *X, double h, int grav) { kaos_ena_push(X, h, grav); }


/* ---- grid set-up ------------------------------------------------------------ */

/* ConfigGridGeneratorSub, ConfigGridGeneratorSubB and VelGridGeneratorSub for
 * one tube, loop for loop over the set-up cells Qind = 1..NqUB+2
 * (_0_ConfigGridGenerator.f90:299-547, _0_1_1_ConfigGridGeneratorB.f90:44-189),
 * the re-index onto loop cells Q = Qind - 1 (_1_SimParameterization.f90:1156-1179)
 * and the velocity axes (_0_0_VelGridGenerator.f90:59-857).  Outputs:
 * cfg[13][nq] = qGL, qGH, qGC, hqC, dqC, d3xC, rGC, cos(thetaGC), ellGC, Ts0,
 * Te0, nsC, nsnormCNeut (truncated); counts[nq] = nsnormC0 (truncated);
 * vel = for the axes V||, Vperp (2-D), Vperp1, Vperp2, Vp, Vq, Vphi (ncell[a]
 * cells, 0 = not generated) [GL | GH | GC | d]; d3v (2-D [Vperp][V||] or 3-D
 * [V1][V2][V||]); hh, d33 [Q][ip][iq][iphi] from set-up cell Q's theta.
 * Returns 0, or -1 for a count that is not finite. */
int orc_generate_grid(int64_t nq_ub, double qGA, double qGB, double L, double ms, double ti, double te, double ns0,
                      double nsnormfac, int vperp3d, int symvpar, int cx, int64_t nvpar_gp, int64_t nvperp_gp,
                      double fpar, double fperp, double tneut, double z0, double n0, double *cfg, int64_t *counts,
                      double *vel, double *d3v, double *hh, double *d33) {
  const int nq = (int)nq_ub, n0c = nq + 2;
  const long long ngp = nq_ub + 3;
  kaos_gg_cell_t *cells = (kaos_gg_cell_t *)malloc(sizeof(kaos_gg_cell_t) * n0c);
  double *icb = (double *)malloc(sizeof(double) * n0c);
  for (int i = 0; i < n0c; ++i) {
    kaos_gg_cell(qGA, qGB, ngp, L, i + 1, &cells[i]);
    icb[i] = kaos_gg_icbbp(&cells[i]);
  }
  /* ICbb = sum(ICbbp(1:Qind)) as a running sum, gC = ICbb - ICbbp(1) */
  double run = icb[0];
  int bad = 0;
  for (int Q = 0; Q < nq; ++Q) {
    const kaos_gg_cell_t *c = &cells[Q + 1];
    run = run + icb[Q + 1];
    const double ns = kaos_gg_ns(run - icb[0], ms, ti, te, ns0, qGA);
    const double x = (ns / nsnormfac) * c->d3x;
    const double xn = cx ? (kaos_gg_ns_neut(c->r, ms, tneut, z0, n0, qGA) / nsnormfac) * c->d3x : 0.0;
    const double v[13] = {c->qGL, c->qGH, c->qGC, c->hq, c->dq, c->d3x, c->r, c->cth, c->ell,
                          (1.0 / 3.0) * ti + (2.0 / 3.0) * ti, te, ns, trunc(xn)};
    for (int k = 0; k < 13; ++k) cfg[(int64_t)k * nq + Q] = v[k];
    if (!(x > -9.2e18 && x < 9.2e18)) bad = 1;
    counts[Q] = bad ? -1 : (int64_t)trunc(x);
  }
  /* velocity axes */
  const double Npar = (double)nvpar_gp / 2.0, Nperp = (double)nvperp_gp / 2.0;
  const double sp = kaos_thermal_speed(ti, ms), sn = cx ? kaos_thermal_speed(tneut, ms) : 0.0;
  const int south = qGA <= 0.0;
  const int nc[7] = {(int)nvpar_gp - 2, vperp3d ? 0 : (int)nvperp_gp - 2, vperp3d ? (int)nvperp_gp - 2 : 0,
                     vperp3d ? (int)nvperp_gp - 2 : 0, cx ? (int)nvperp_gp - 2 : 0, cx ? (int)nvpar_gp - 2 : 0,
                     cx ? (int)nvperp_gp - 2 : 0};
  int off[7], tot = 0;
  for (int a = 0; a < 7; ++a) { off[a] = tot; tot += 4 * nc[a]; }
  double P[520];
  for (int a = 0; a < 7; ++a) {
    const int n = nc[a];
    for (int i = 1; i <= n + 1; ++i) {
      double x;
      if (a == 0) x = symvpar ? kaos_gg_vsym(Npar, fpar * sp, i) : kaos_gg_vlin((double)nvpar_gp, -4.0 * sp, fpar * sp, south, i);
      else if (a <= 3) x = kaos_gg_vsym(Nperp, fperp * sp, i);
      else if (a == 5) x = symvpar ? kaos_gg_vena(Npar, fpar * sn, i) : kaos_gg_vlin((double)nvpar_gp, -4.0 * sn, fpar * sn, south, i);
      else x = kaos_gg_vena(Nperp, fperp * sn, i);
      P[i] = x;
    }
    double *b = vel + off[a];
    for (int i = 0; i < n; ++i) {
      b[i] = P[i + 1];
      b[n + i] = P[i + 2];
      b[2 * n + i] = (P[i + 1] + P[i + 2]) / 2.0;
      if (a == 0 && b[2 * n + i] == 0.0) b[2 * n + i] = 1e-15;
      b[3 * n + i] = fabs(P[i + 2] - P[i + 1]);
    }
  }
  const int np = nc[0];
  if (vperp3d) {
    const double *c1 = vel + off[2], *c2 = vel + off[3], *cp = vel + off[0];
    const int n1 = nc[2], n2 = nc[3];
    for (int i1 = 0; i1 < n1; ++i1)
      for (int i2 = 0; i2 < n2; ++i2)
        for (int ip = 0; ip < np; ++ip)
          d3v[((int64_t)i1 * n2 + i2) * np + ip] = (((1.0 * 1.0) * 1.0 * c1[3 * n1 + i1]) * c2[3 * n2 + i2]) * cp[3 * np + ip];
  } else {
    const double *c1 = vel + off[1], *cp = vel + off[0];
    const int n1 = nc[1];
    for (int i1 = 0; i1 < n1; ++i1)
      for (int ip = 0; ip < np; ++ip) {
        double d = ((((1.0 * 1.0) * fabs(c1[2 * n1 + i1])) * c1[3 * n1 + i1]) * cp[3 * np + ip]) * (2.0 * KAOS_PI);
        if (d == 0.0) d = 1e-9;
        d3v[(int64_t)i1 * np + ip] = d;
      }
  }
  if (cx) {
    const double *vp = vel + off[4], *vq = vel + off[5], *vf = vel + off[6];
    const int a = nc[4], b = nc[5], f = nc[6];
    int64_t k = 0;
    for (int Q = 0; Q < nq; ++Q)
      for (int ip = 0; ip < a; ++ip)
        for (int iq = 0; iq < b; ++iq)
          for (int iphi = 0; iphi < f; ++iphi, ++k)
            kaos_gg_ena_metric(cells[Q].cth, cells[Q].sth, cells[Q].ell, vp[2 * a + ip], vq[2 * b + iq], vp[3 * a + ip],
                               vq[3 * b + iq], vf[3 * f + iphi], &hh[k], &d33[k]);
  }
  free(cells);
  free(icb);
  return bad ? -1 : 0;
}

/* Shared-header pieces of the set-up, for the pins. */
double orc_gg_qpoint(double qGA, double qGB, int64_t ngp, int64_t i) { return kaos_gg_qpoint(qGA, qGB, ngp, i); }
double orc_sin(double x) { return kaos_sin(x); }
void orc_gg_ena_metric(double cth, double sth, double ell, double vp, double vq, double dvp, double dvq, double dvphi,
                       double *o2) {
  kaos_gg_ena_metric(cth, sth, ell, vp, vq, dvp, dvq, dvphi, &o2[0], &o2[1]);
}
