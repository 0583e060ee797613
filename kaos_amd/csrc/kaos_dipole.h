/*
 * kaos_dipole.h -- scalar dipole-geometry, RK4 and random-variate arithmetic of
 * the KAOS kinetic hot path, compiled for both the gfx950 kernels (hipcc) and
 * the host (gcc/clang).
 *
 * Every function here evaluates a fixed sequence of IEEE-754 basic operations
 * (+ - * / sqrt, compares, frexp/ldexp) with floating-point contraction
 * disabled by the build (-ffp-contract=off on both compilers).  Those
 * operations are correctly rounded on gfx950 and on x86-64 SSE2, so a kernel
 * and a host caller of the same function on the same inputs produce the same
 * bits.
 *
 * Where the reference evaluates a transcendental of a transcendental, the
 * exact algebraic value is used instead (all equal in real arithmetic):
 *   cos(acos(z/r))            -> z/r                  (_KineticMainParams.f90:579-582)
 *   sin(acos(z/r))            -> rho/r, rho = sqrt(x^2+y^2)  (theta in [0,pi])
 *   sin(asin(sqrt(w)))        -> sqrt(w)              (_8_1_4_RK4DipolePolynomialSolver.f90:89-90)
 *   cos(asin(sqrt(w)))        -> sqrt(1-w)
 *   cos(pi - t)               -> -cos(t)              (_8_1_4_RK4DipolePolynomialSolver.f90:101-105)
 *   Re{2 i sqrt(e) cos((acos(D/(i e^1.5)) + 4 pi)/3)}
 *                             -> -2 sqrt(e) sinh(asinh(D/e^1.5)/3)
 *                                                     (_8_1_4_RK4DipolePolynomialSolver.f90:58-67)
 * and w = sinh(asinh(y)/3) is obtained as the real root of the triple-angle
 * identity 4 w^3 + 3 w - y = 0 by a fixed number of Newton steps.
 */
#ifndef KAOS_DIPOLE_H
#define KAOS_DIPOLE_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define KAOS_HD __host__ __device__ static inline
#else
#define KAOS_HD static inline
#endif

/* Physical constants, _ModelInput.f90:35-43 (same decimal literals). */
#define KAOS_RE   6.38e6          /* RE= 638d4   [m]               */
#define KAOS_RE2  4.07044e13      /* RE**2 (exact in binary64)     */
#define KAOS_GG   6.67e-11        /* GG= 667d-13                   */
#define KAOS_ME   5.98e24         /* ME= 598d22  [kg]              */
#define KAOS_MDIP 8.06e15         /* m= 8.06d15  [T m^3]           */
#define KAOS_KB   1.38e-23        /* kB= 138d-25                   */
#define KAOS_PI   3.14159265358979311600   /* 4*atan(1d0) in binary64 */
#define KAOS_TWOPI 6.28318530717958623200
#define KAOS_ZERO_CUT 1e-6        /* _8_1_4_RK4DipolePolynomialSolver.f90:115-125 */

/* ------------------------------------------------------------------------- */
/* Resolvent-cubic root: w = sinh(asinh(y)/3), y >= 0, i.e. the real root of
 * 4 w^3 + 3 w - y = 0.  Seeds: Taylor y/3 - 4y^3/81 below 1, a frexp-based
 * cube-root estimate of (y/4)^(1/3) above; then 7 Newton steps
 * w <- (8 w^3 + y)/(12 w^2 + 3).                                            */
KAOS_HD double kaos_sinh_asinh_third(double y) {
  double w;
  int nit;
  if (y < 1.0) {
    double y2 = y * y;
    w = y * ((1.0 / 3.0) - (4.0 / 81.0) * y2);
    nit = 4;
  } else {
    int e;
    double m = frexp(0.25 * y, &e);             /* 0.25y = m 2^e, m in [0.5,1) */
    int e3 = (e >= 0) ? (e / 3) : -((-e + 2) / 3);
    int rem = e - 3 * e3;                       /* 0,1,2 */
    double mm = ldexp(m, rem);                  /* in [0.5, 4) */
    double cr = 0.5 + mm * (0.55 - mm * 0.05);  /* crude cbrt on [0.5,4) */
    w = ldexp(cr, e3);
    if (w < 0.3) w = 0.3;
    nit = 7;
  }
  for (int it = 0; it < nit; ++it) {
    double w2 = w * w;
    w = (8.0 * (w2 * w) + y) / (12.0 * w2 + 3.0);
  }
  return w;
}

/* RK4DipolePolynomialSolverSub (_8_1_4_RK4DipolePolynomialSolver.f90:26-88):
 * the positive real root gamma of gamma^4 + C gamma + D = 0 with
 * C = 1/(p q^2), D = -1/q^2 (Ferrari via the resolvent cubic), returned as
 * r = gamma RE.  The final subtraction (pii - mu) is the reference's
 * cancellation-prone step and is kept as is.                               */
KAOS_HD double kaos_root_r(double q, double p) {
  double iq2 = 1.0 / (q * q);
  double C = iq2 * (1.0 / p);                   /* CpRK4 = 1/(p q^2) */
  double eps = iq2 * (4.0 / 3.0);               /* EpsilonRK4 = Abb/3, Abb = -4 Dp */
  double delta = (C * C) * 0.5;                 /* DeltaRK4 = -Bbb/2, Bbb = -Cp^2 */
  double se = sqrt(eps);
  double y = delta / (eps * se);
  double w = kaos_sinh_asinh_third(y);
  double sigma = (2.0 * se) * w;                /* sigmaRK4 */
  double mu = sqrt(sigma);                      /* muRK4 */
  double pii = sqrt((2.0 * C) / mu - mu * mu);  /* piiRK4 */
  double gamma = 0.5 * (pii - mu);              /* gamma3RK4 */
  return gamma * KAOS_RE;                       /* r3RK4 */
}

/* Signed radius of the field-line point with dipole coordinate q. */
KAOS_HD double kaos_rs_of_q(double q, double p) {
  double r = kaos_root_r(q, p);
  return (q <= 0.0) ? -r : r;
}

/* ------------------------------------------------------------------------- */
/* Cell searches.  Contiguous edges (qGL(Q+1) == qGH(Q), VGL(i+1) == VGH(i))
 * make the reference's first-match linear scans equivalent to "first cell
 * whose high edge is not passed".                                          */

/* Config cell of q (_8_1_1_KineticUpdateA.f90:221-273): 1..nq, 0 = below the
 * lower boundary, -1 = beyond the upper boundary.  desc != 0 when q decreases
 * with the cell index (qGL(1) > 0).  qGH is the nq high edges, qGL0 the low
 * edge of cell 1.  NaN q is reported as 0.                                 */
KAOS_HD int kaos_find_qcell(double q, const double* qGH, double qGL0, double qGHlast,
                            int nq, int desc) {
  if (desc) {
    if (!(q <= qGL0)) return 0;          /* qGL(NqLB) < q, or NaN */
    if (qGHlast > q) return -1;
    int lo = 0, hi = nq - 1;             /* first i with q >= qGH[i] */
    while (lo < hi) { int mid = (lo + hi) >> 1; if (q >= qGH[mid]) hi = mid; else lo = mid + 1; }
    return lo + 1;
  } else {
    if (!(q >= qGL0)) return 0;          /* qGL(NqLB) > q, or NaN */
    if (qGHlast < q) return -1;
    int lo = 0, hi = nq - 1;             /* first i with q <= qGH[i] */
    while (lo < hi) { int mid = (lo + hi) >> 1; if (q <= qGH[mid]) hi = mid; else lo = mid + 1; }
    return lo + 1;
  }
}

/* Velocity cell of v on ascending contiguous edges (cell 1 is [GL,GH], the
 * others (GL,GH]; _8_1_1_KineticUpdateA.f90:1087-1127).  1..n, 0 if off grid. */
KAOS_HD int kaos_find_vcell(double v, const double* GH, double GL0, int n) {
  if (!(v >= GL0) || !(v <= GH[n - 1])) return 0;
  int lo = 0, hi = n - 1;
  while (lo < hi) { int mid = (lo + hi) >> 1; if (v <= GH[mid]) hi = mid; else lo = mid + 1; }
  return lo + 1;
}

/* The same two searches started from a hint (the ion's cell of the previous
 * step, 1-based; ions move less than a cell per step): a walk to the first
 * cell whose high edge is not passed, identical in result to the bisection. */
KAOS_HD int kaos_find_qcell_from(double q, const double* qGH, double qGL0, double qGHlast,
                                 int nq, int desc, int hint) {
  int i = hint - 1;
  if (i < 0) i = 0;
  if (i > nq - 1) i = nq - 1;
  if (desc) {
    if (!(q <= qGL0)) return 0;
    if (qGHlast > q) return -1;
    while (i > 0 && q >= qGH[i - 1]) --i;
    while (!(q >= qGH[i])) ++i;
  } else {
    if (!(q >= qGL0)) return 0;
    if (qGHlast < q) return -1;
    while (i > 0 && q <= qGH[i - 1]) --i;
    while (!(q <= qGH[i])) ++i;
  }
  return i + 1;
}

/* Velocity cell by a walk from a hint cell (1-based; 0 or out of range:
 * bisection), the same result as kaos_find_vcell on ascending contiguous
 * edges: the first cell whose high edge is not passed.                    */
KAOS_HD int kaos_find_vcell_from(double v, const double* GH, double GL0, int n, int hint) {
  if (!(v >= GL0) || !(v <= GH[n - 1])) return 0;
  if (hint < 1 || hint > n) return kaos_find_vcell(v, GH, GL0, n);
  int i = hint - 1;
  while (i > 0 && v <= GH[i - 1]) --i;
  while (!(v <= GH[i])) ++i;
  return i + 1;
}

/* Velocity cell from a uniform-spacing guess (scale = n/(GH[n-1]-GL0)). */
KAOS_HD int kaos_find_vcell_guess(double v, const double* GH, double GL0, int n, double scale) {
  if (!(v >= GL0) || !(v <= GH[n - 1])) return 0;
  int i = (int)((v - GL0) * scale);
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  while (i > 0 && v <= GH[i - 1]) --i;
  while (!(v <= GH[i])) ++i;
  return i + 1;
}

/* ------------------------------------------------------------------------- */
/* Field interpolation along the tube (_8_7_AmbipolarEfield.f90:1666-1860 for
 * q increasing with the cell index, 2067-2230 for q decreasing;
 * _8_7_2_Gravfield.f90:132-300).  Returns 1 and the interpolated value when q
 * falls in [qGL(LB), qGH(UB)]; 0 (value untouched) otherwise.             */
/* Interpolation with a walk from a hint cell (1-based) instead of bisection;
 * E2 (may be NULL) is interpolated on the same interval into *out2.        */
KAOS_HD int kaos_interp2_from(double q, const double* qGC, const double* E, const double* E2, int nq,
                              double qGL0, double qGHlast, int desc, int hint, double* out,
                              double* out2) {
  int ub = nq - 1;
  int lo = hint - 1;
  if (lo < 1) lo = 1;
  if (lo > ub) lo = ub;
  if (desc) {
    if (!(q <= qGL0) || !(q >= qGHlast)) return 0;
    if (q >= qGC[0]) { *out = E[0]; if (E2) *out2 = E2[0]; return 1; }
    if (q <= qGC[ub]) { *out = E[ub]; if (E2) *out2 = E2[ub]; return 1; }
    while (lo > 1 && q >= qGC[lo - 1]) --lo;
    while (!(q >= qGC[lo])) ++lo;
  } else {
    if (!(q >= qGL0) || !(q <= qGHlast)) return 0;
    if (q <= qGC[0]) { *out = E[0]; if (E2) *out2 = E2[0]; return 1; }
    if (q >= qGC[ub]) { *out = E[ub]; if (E2) *out2 = E2[ub]; return 1; }
    while (lo > 1 && q <= qGC[lo - 1]) --lo;
    while (!(q <= qGC[lo])) ++lo;
  }
  double x1 = qGC[lo - 1], x2 = qGC[lo];
  double f = (q - x1) / (x2 - x1);
  *out = E[lo - 1] + (f * (E[lo] - E[lo - 1]));
  if (E2) *out2 = E2[lo - 1] + (f * (E2[lo] - E2[lo - 1]));
  return 1;
}

KAOS_HD int kaos_interp_cells(double q, const double* qGC, const double* E, int nq,
                              double qGL0, double qGHlast, int desc, double* out) {
  int ub = nq - 1;
  if (desc) {
    if (!(q <= qGL0) || !(q >= qGHlast)) return 0;
    if (q >= qGC[0]) { *out = E[0]; return 1; }          /* LB half cell */
    if (q <= qGC[ub]) { *out = E[ub]; return 1; }        /* UB half cell */
    int lo = 1, hi = ub;                 /* first Q with q >= qGC[Q] */
    while (lo < hi) { int mid = (lo + hi) >> 1; if (q >= qGC[mid]) hi = mid; else lo = mid + 1; }
    double x1 = qGC[lo - 1], x2 = qGC[lo], y1 = E[lo - 1], y2 = E[lo];
    double f = (q - x1) / (x2 - x1);
    *out = y1 + (f * (y2 - y1));
    return 1;
  } else {
    if (!(q >= qGL0) || !(q <= qGHlast)) return 0;
    if (q <= qGC[0]) { *out = E[0]; return 1; }
    if (q >= qGC[ub]) { *out = E[ub]; return 1; }
    int lo = 1, hi = ub;                 /* first Q with q <= qGC[Q] */
    while (lo < hi) { int mid = (lo + hi) >> 1; if (q <= qGC[mid]) hi = mid; else lo = mid + 1; }
    double x1 = qGC[lo - 1], x2 = qGC[lo], y1 = E[lo - 1], y2 = E[lo];
    double f = (q - x1) / (x2 - x1);
    *out = y1 + (f * (y2 - y1));
    return 1;
  }
}

/* ------------------------------------------------------------------------- */
/* One RK4 push of one ion (_8_1_KineticRK4Update.f90:29-314 with
 * KineticUpdateA/B/C).  Forces: gravity and ambipolar field only (mirror and
 * parallel-potential flags are off on this path).  The parallel acceleration
 * of each stage is (sG |AG| + sE AEA) and the Cartesian acceleration is that
 * times b-hat; the reference projects each force separately and sums, which
 * is the same in real arithmetic.
 *
 * An ion stays in the meridional plane of its tube (the reference resets phi
 * to the tube longitude every step, _8_1_KineticRK4Update.f90:165-166), so
 * the push is carried out on (rho, z) = (sqrt(x^2+y^2), z), with
 * (x, y) = rho (cos phi, sin phi); and the state at a step boundary is the
 * signed radius rs of the root solve plus V|| (the reference's state is
 * x = pos(rs), V = V|| b-hat(x), _8_1_KineticRK4Update.f90:170-314).        */

typedef struct {
  double h;            /* time step SpecieT%hT */
  double p;            /* L-shell pGCGT */
  double ip;           /* 1/p */
  double irep;         /* 1/(RE p) */
  double cphi, sphi;   /* cos/sin of the tube longitude phiGCGT */
  int grav;            /* GRAVflagT */
  int eamb;            /* EAMBflagT */
  int eambsign;        /* EAMBSIGNflagT */
} kaos_push_par_t;

KAOS_HD kaos_push_par_t kaos_push_par(double h, double p, double cphi, double sphi, int grav,
                                      int eamb, int eambsign) {
  kaos_push_par_t P;
  P.h = h; P.p = p; P.ip = 1.0 / p; P.irep = 1.0 / (KAOS_RE * p);
  P.cphi = cphi; P.sphi = sphi;
  P.grav = grav; P.eamb = eamb; P.eambsign = eambsign;
  return P;
}

typedef struct {
  double rsN;            /* signed radius of the new point                 */
  double vparN;          /* VparNp: parallel velocity of the new state     */
  double vpar;           /* Vpar(j): parallel velocity of the input state  */
  double qk1;            /* q of the input state (cell search input)       */
  double qk4;            /* q at the full-step predictor (field coordinate) */
  double zC, u2C;        /* predictor z and rho^2 + z^2: cos(thetak4) = zC / sqrt(u2C) */
} kaos_push_out_t;

/* Point of the field line with signed radius rs: r, w = r/(RE p) = sin^2,
 * cos, sin and (rho, z), with the reference's zeroing of tiny coordinates
 * (_8_1_4_RK4DipolePolynomialSolver.f90:111-125).                         */
typedef struct { double r, w, c, s, rho, z; } kaos_line_t;

KAOS_HD kaos_line_t kaos_line_i(double rs, double irep) {
  kaos_line_t L;
  L.r = fabs(rs);
  L.w = L.r * irep;
  L.s = sqrt(L.w);
  L.c = sqrt(1.0 - L.w);
  if (rs < 0.0) L.c = -L.c;
  L.rho = L.r * L.s;
  L.z = L.r * L.c;
  if (fabs(L.z) < KAOS_ZERO_CUT) L.z = 0.0;
  return L;
}

KAOS_HD kaos_line_t kaos_line(double rs, double p) { return kaos_line_i(rs, 1.0 / (KAOS_RE * p)); }

/* q and b-hat (rho, z components) at a point of the meridional plane. */
typedef struct { double q, bR, bZ; } kaos_qb_t;

/* On the line: b-hat = (3 c s, 3 c^2 - 1)/sqrt(1 + 3 c^2) with c^2 = 1 - w
 * (_KineticMainParams.f90:614, 674-779), q = RE^2 c / r^2 (:593).           */
KAOS_HD kaos_qb_t kaos_stageA(kaos_line_t L) {
  kaos_qb_t g;
  g.q = (KAOS_RE2 * L.c) / (L.r * L.r);
  double isl = 1.0 / sqrt(4.0 - 3.0 * L.w);
  g.bR = (3.0 * (L.c * L.s)) * isl;
  g.bZ = (2.0 - 3.0 * L.w) * isl;
  return g;
}

/* Off the line, at (rho, z) with u2 = rho^2 + z^2 = r^2: c = z/r, s = rho/r,
 * so b-hat = (3 z rho, 2 z^2 - rho^2) / sqrt(u2 (u2 + 3 z^2)): one square
 * root and one division.                                                    */
KAOS_HD void kaos_bhat_pt(double rho, double z, double* bR, double* bZ, double* u2out) {
  double z2 = z * z;
  double u2 = rho * rho + z2;
  double d = 1.0 / sqrt(u2 * (u2 + 3.0 * z2));
  *bR = (3.0 * (z * rho)) * d;
  *bZ = (2.0 * z2 - rho * rho) * d;
  *u2out = u2;
}

/* Cartesian point and velocity V|| b-hat of a line state (for output). */
KAOS_HD void kaos_cartesian(double rs, double vs, double p, double cphi, double sphi,
                            double* x, double* y, double* z, double* Vx, double* Vy, double* Vz) {
  kaos_line_t L = kaos_line(rs, p);
  double xx = L.rho * cphi, yy = L.rho * sphi;
  if (fabs(xx) < KAOS_ZERO_CUT) xx = 0.0;
  if (fabs(yy) < KAOS_ZERO_CUT) yy = 0.0;
  *x = xx; *y = yy; *z = L.z;
  kaos_qb_t g = kaos_stageA(L);
  *Vx = (vs * g.bR) * cphi; *Vy = (vs * g.bR) * sphi; *Vz = vs * g.bZ;
}

/* Parallel acceleration of one stage; only the sign of q enters
 * (sign conventions of _8_1_1_KineticUpdateA.f90:480-484, 649-658).       */
KAOS_HD double kaos_apar(double q, double AG, double AEA, const kaos_push_par_t* P, int with_grav) {
  double a = 0.0;
  if (P->grav && with_grav) a = ((q <= 0.0) ? 1.0 : -1.0) * fabs(AG);        /* AGmagS, KUA:480-484 */
  if (P->eamb) {
    if (P->eambsign) a = a + AEA;                                             /* AEAmagS=1, KUA:656-658 */
    else a = a + ((q <= 0.0) ? -1.0 : 1.0) * fabs(AEA);                       /* KUA:649-655 */
  }
  return a;
}

/* Positive root of the dipole quartic (_8_1_4_RK4DipolePolynomialSolver.f90:30-33,
 * gamma^4 + gamma/(p q^2) - 1/q^2 = 0) for the point (R, Z) of the meridional
 * plane, q^2 = RE^4 Z^2 / (R^2 + Z^2)^3 = A / B.  Multiplied by B q^2 the
 * quartic is f(g) = A g^4 + B (g/p - 1); f is increasing and convex for
 * g > 0, so Newton's method converges monotonically from any g0 > 0,
 * quadratically with relative error e -> <= 1.5 e^2, one division per step.
 * Steps continue while the last correction exceeds 1e-9 g; from a start
 * within 1e-9 of the root one step reaches it to rounding (1.5e-18).       */
KAOS_HD double kaos_quartic_newton(double A, double B, double ip, double g0) {
  double g = g0, dg;
  int it = 0;
  do {
    double g2 = g * g;
    double f = (A * g2) * g2 + B * (g * ip - 1.0);
    double fp = (4.0 * A) * (g2 * g) + B * ip;
    dg = f / fp;
    g = g - dg;
    ++it;
  } while (fabs(dg) > 1e-9 * g && it < 64);
  return g;
}

/* The push, given the stage-A geometry (L, gA) and the input cell qc of the
 * ion (used only for the stage-4 gravity gate, _8_1_3_KineticUpdateC.f90:345-346). */
KAOS_HD void kaos_push_ion(kaos_line_t L, kaos_qb_t gA, int qc, double vs, double AG, double AEA,
                           const kaos_push_par_t* P, kaos_push_out_t* o) {
  const double h = P->h, h2 = P->h / 2.0, hsq2 = (P->h * P->h) / 2.0, h6 = P->h / 6.0;
  double ap1 = kaos_apar(gA.q, AG, AEA, P, 1);
  double a1R = ap1 * gA.bR, a1Z = ap1 * gA.bZ;
  /* V = V|| b-hat; Vpar(j) = b-hat . V (KUA:978-980), VR = Vpar b-hat (KUA:1327-1334) */
  double vpar = gA.bR * (vs * gA.bR) + gA.bZ * (vs * gA.bZ);
  double VRR = vpar * gA.bR, VRZ = vpar * gA.bZ;
  /* half-step position and stage B (KineticUpdateBSub): sign(q) = sign(z) */
  double b2R, b2Z, u2B;
  double zB = L.z + h2 * VRZ;
  kaos_bhat_pt(L.rho + h2 * VRR, zB, &b2R, &b2Z, &u2B);
  double ap2 = kaos_apar(zB, AG, AEA, P, 1);
  double a2R = ap2 * b2R, a2Z = ap2 * b2Z;
  /* full-step predictor and stage C (KineticUpdateCSub): q4 = RE^2 z / r^3 */
  double b4R, b4Z, u24;
  double zC = (L.z + h * VRZ) + hsq2 * a1Z;
  kaos_bhat_pt((L.rho + h * VRR) + hsq2 * a1R, zC, &b4R, &b4Z, &u24);
  double q4 = (KAOS_RE2 * zC) / (u24 * sqrt(u24));
  double ap4 = kaos_apar(q4, AG, AEA, P, (qc != 0 && qc != -1));
  double a4R = ap4 * b4R, a4Z = ap4 * b4Z;
  /* velocity: k1 + 2 k2 + 2 k3 + k4 with k2 = k3 (_8_1_KineticRK4Update.f90:85-90) */
  double VNR = VRR + h6 * (((a1R + 2.0 * a2R) + 2.0 * a2R) + a4R);
  double VNZ = VRZ + h6 * (((a1Z + 2.0 * a2Z) + 2.0 * a2Z) + a4Z);
  /* position (_8_1_KineticRK4Update.f90:124-156) */
  double k2R = VRR + h2 * a1R, k2Z = VRZ + h2 * a1Z;
  double k4R = VRR + h * a2R, k4Z = VRZ + h * a2Z;
  double RNp = L.rho + h6 * (((VRR + 2.0 * k2R) + 2.0 * k2R) + k4R);
  double ZNp = L.z + h6 * (((VRZ + 2.0 * k2Z) + 2.0 * k2Z) + k4Z);
  /* re-project onto the L-shell (:161-176): the point of the line with the
   * same q, hemisphere from sign(Z); Newton starts from the first-order
   * radius r + r-hat . (dR, dZ) (relative error ~ |dx|^2 / r^2 ~ 1e-10) */
  double Z2 = ZNp * ZNp, u2N = RNp * RNp + Z2;
  double g0 = (L.r + (L.s * (RNp - L.rho) + L.c * (ZNp - L.z))) * (1.0 / KAOS_RE);
  double rN = kaos_quartic_newton((KAOS_RE2 * KAOS_RE2) * Z2, (u2N * u2N) * u2N, P->ip, g0) * KAOS_RE;
  double rsN = (ZNp <= 0.0) ? -rN : rN;
  /* re-align the velocity with b-hat at the new point (:273-314):
   * c s = sign * sqrt(w (1 - w)), 3 c^2 - 1 = 2 - 3 w */
  double wN = rN * P->irep;
  double csN = sqrt(wN * (1.0 - wN));
  if (rsN < 0.0) csN = -csN;
  double islN = 1.0 / sqrt(4.0 - 3.0 * wN);
  o->rsN = rsN;
  o->vparN = ((3.0 * csN) * islN) * VNR + ((2.0 - 3.0 * wN) * islN) * VNZ;
  o->vpar = vpar;
  o->qk1 = gA.q;
  o->qk4 = q4;
  o->zC = zC;
  o->u2C = u24;
}

/* ------------------------------------------------------------------------- */
/* Random variates.                                                          */

/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter ctr, key key. */
KAOS_HD void kaos_mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  *lo = a * b;
  *hi = __umul64hi(a, b);
#else
  unsigned __int128 pr = (unsigned __int128)a * (unsigned __int128)b;
  *lo = (uint64_t)pr;
  *hi = (uint64_t)(pr >> 64);
#endif
}

KAOS_HD void kaos_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
    uint64_t hi0, lo0, hi1, lo1;
    kaos_mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
    kaos_mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 64 random bits -> uniform double in (0, 1]. */
KAOS_HD double kaos_u01(uint64_t b) {
  return (double)((b >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
}

/* Natural log for finite x > 0: frexp reduction to [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh((m-1)/(m+1)) by its odd series, plus e ln2 in two parts. */
KAOS_HD double kaos_log(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  double f = (m - 1.0) / (m + 1.0);
  double s = f * f;
  double pp = 1.0 / 25.0;
  pp = pp * s + 1.0 / 23.0; pp = pp * s + 1.0 / 21.0; pp = pp * s + 1.0 / 19.0;
  pp = pp * s + 1.0 / 17.0; pp = pp * s + 1.0 / 15.0; pp = pp * s + 1.0 / 13.0;
  pp = pp * s + 1.0 / 11.0; pp = pp * s + 1.0 / 9.0;  pp = pp * s + 1.0 / 7.0;
  pp = pp * s + 1.0 / 5.0;  pp = pp * s + 1.0 / 3.0;
  double f2 = 2.0 * f;
  double lm = f2 + (f2 * s) * pp;
  double de = (double)e;
  return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lm);
}

/* cos(2 pi u) for u in [0, 1]: octant reduction (exact subtractions) and the
 * Taylor series of cos/sin on [-pi/4, pi/4].                               */
KAOS_HD double kaos_cos_poly(double x) {
  double s = x * x;
  double p = -1.0 / 2432902008176640000.0;          /* -1/20! */
  p = p * s + 1.0 / 6402373705728000.0;             /* 1/18! */
  p = p * s - 1.0 / 20922789888000.0;               /* 1/16! */
  p = p * s + 1.0 / 87178291200.0;
  p = p * s - 1.0 / 479001600.0;
  p = p * s + 1.0 / 3628800.0;
  p = p * s - 1.0 / 40320.0;
  p = p * s + 1.0 / 720.0;
  p = p * s - 1.0 / 24.0;
  p = p * s + 0.5;
  return 1.0 - s * p;    /* 1 - x^2/2 + x^4/24 - ... */
}
KAOS_HD double kaos_sin_poly(double x) {
  double s = x * x;
  double p = 1.0 / 121645100408832000.0;            /* 1/19! */
  p = p * s - 1.0 / 355687428096000.0;              /* 1/17! */
  p = p * s + 1.0 / 1307674368000.0;
  p = p * s - 1.0 / 6227020800.0;
  p = p * s + 1.0 / 39916800.0;
  p = p * s - 1.0 / 362880.0;
  p = p * s + 1.0 / 5040.0;
  p = p * s - 1.0 / 120.0;
  p = p * s + 1.0 / 6.0;
  return x - (x * s) * p;
}
/* sin(2 pi u) for u in [0, 1], with the same exact octant reductions. */
KAOS_HD double kaos_sin2pi(double u) {
  double t = u, sg = 1.0;
  if (t > 0.5) { t = t - 0.5; sg = -1.0; }     /* sin(x + pi) = -sin x   */
  if (t > 0.25) t = 0.5 - t;                    /* sin(pi - x) = sin x    */
  double r = (t <= 0.125) ? kaos_sin_poly(KAOS_TWOPI * t) : kaos_cos_poly(KAOS_TWOPI * (0.25 - t));
  return sg * r;
}
KAOS_HD double kaos_cos2pi(double u) {
  double t = u;
  if (t > 0.5) t = 1.0 - t;                     /* cos symmetric about 1/2 */
  if (t <= 0.125) return kaos_cos_poly(KAOS_TWOPI * t);
  if (t <= 0.375) return kaos_sin_poly(KAOS_TWOPI * (0.25 - t));
  return -kaos_cos_poly(KAOS_TWOPI * (0.5 - t));
}

/* The seeded uniform stream of one injected ion: six uniforms (Vx, Vy, Vz
 * Box-Muller pairs, _6_1_MBVelocityDistribution.f90:34-47) from two Philox
 * blocks keyed by (seed, "KAOS") with counter (ion k of the step, step n,
 * tube, block).                                                            */
KAOS_HD void kaos_inject_uniforms(uint64_t seed, int64_t step, int tube, int64_t k, double u[6]) {
  uint64_t key[2] = { seed, 0x4B414F53ULL };
  uint64_t ctr[4] = { (uint64_t)k, (uint64_t)step, (uint64_t)tube, 0x494E4A00ULL };
  uint64_t o[4];
  kaos_philox4x64(ctr, key, o);
  u[0] = kaos_u01(o[0]); u[1] = kaos_u01(o[1]); u[2] = kaos_u01(o[2]); u[3] = kaos_u01(o[3]);
  ctr[3] = 0x494E4A01ULL;
  kaos_philox4x64(ctr, key, o);
  u[4] = kaos_u01(o[0]); u[5] = kaos_u01(o[1]);
}

/* Box-Muller Gaussian (_3_1_GaussianRNG.f90:31-33). */
KAOS_HD double kaos_gauss(double u1, double u2) {
  return sqrt(-2.0 * kaos_log(u1)) * kaos_cos2pi(u2);
}

/* Maxwell-Boltzmann injection velocity (_6_1_MBVelocityDistribution.f90:29-108)
 * at the field-line point with geometry g: projection of an isotropic
 * Cartesian Maxwellian onto (Vp, Vq, Vphi); Vperp1 = Vp, Vperp2 = Vphi,
 * Vperp = |(Vperp1, Vperp2)|, Vpar = Vq.                                    */
KAOS_HD void kaos_mb_project(const double u[6], double sigma, double cphi, double sphi,
                             kaos_line_t g, double* vpar, double* vperp1,
                             double* vperp2, double* vperp) {
  double Vx = sigma * kaos_gauss(u[0], u[1]);
  double Vy = sigma * kaos_gauss(u[2], u[3]);
  double Vz = sigma * kaos_gauss(u[4], u[5]);
  double ell = 1.0 + 3.0 * (g.c * g.c);
  double sl = sqrt(ell);
  double h = Vx * cphi + Vy * sphi;
  double Vp = ((1.0 - 3.0 * (g.c * g.c)) / sl) * h + (((3.0 * Vz) * g.c) * g.s) / sl;
  double Vq = (((3.0 * g.c) * g.s) / sl) * h + (Vz * (3.0 * (g.c * g.c) - 1.0)) / sl;
  double Vphi = -Vx * sphi + Vy * cphi;
  *vperp1 = Vp;
  *vperp2 = Vphi;
  *vperp = fabs(sqrt(Vp * Vp + Vphi * Vphi));
  *vpar = Vq;
}

/* ------------------------------------------------------------------------- */
/* Initial ions (DensityProfileA1Sub, DensityProfileBSub, DipolePolynomial-  */
/* SolverSub, VelocityDistributionSub, InitialConditionsSub).                */

/* Macro-ions of one config cell: (n / nsnormfac) d3x truncated toward zero.
 * The reference assigns the value to nsnormC0, an integer(kind=dp) array
 * (_KineticMainParams.f90:64; _0_1_1_ConfigGridGeneratorB.f90:83-85, and
 * :72-75 for a restart's DensityInputT), which truncates; NsFARRp =
 * nint(nsnormCGT) (_2_1_DensityProfileA1.f90:45) then keeps that integer.
 * -1 when the value is not finite or does not fit in int64.                */
KAOS_HD int64_t kaos_cell_count(double density, double d3x, double nsnormfac) {
  double x = (density / nsnormfac) * d3x;
  if (!(x > -9.2e18 && x < 9.2e18)) return -1;
  return (int64_t)trunc(x);
}

/* Thermal speed sqrt(kB T / m) of MBVelocityDistributionSub
 * (_6_1_MBVelocityDistribution.f90:42), in its order of operations.        */
KAOS_HD double kaos_thermal_speed(double T, double ms) { return sqrt((KAOS_KB * T) / ms); }

/* The seeded uniform stream of initial ion i of a tube (ions numbered cell
 * by cell, as InitialConditionsSub concatenates them,
 * _7_InitialConditions.f90:72-87): six uniforms for the Vx, Vy, Vz
 * Box-Muller pairs of MBVelocityDistributionSub, from two Philox4x64-10
 * blocks keyed by (seed, "KAOS") with counter (i, 0, tube, "INI" block).
 * The "INI" words keep this stream apart from the injection stream's "INJ"
 * blocks.                                                                  */
KAOS_HD void kaos_seed_uniforms(uint64_t seed, int tube, int64_t i, double u[6]) {
  uint64_t key[2] = { seed, 0x4B414F53ULL };
  uint64_t ctr[4] = { (uint64_t)i, 0ULL, (uint64_t)tube, 0x494E4900ULL };
  uint64_t o[4];
  kaos_philox4x64(ctr, key, o);
  u[0] = kaos_u01(o[0]); u[1] = kaos_u01(o[1]); u[2] = kaos_u01(o[2]); u[3] = kaos_u01(o[3]);
  ctr[3] = 0x494E4901ULL;
  kaos_philox4x64(ctr, key, o);
  u[4] = kaos_u01(o[0]); u[5] = kaos_u01(o[1]);
}

/* One initial ion at dipole coordinate q of L-shell p (DensityProfileBSub
 * puts it at its cell centre, _4_DensityProfileB.f90:40-46): signed radius
 * from the root solve (DipolePolynomialSolverSub), Maxwell-Boltzmann
 * velocity of thermal speed sigma at that point from the uniforms u
 * (VelocityDistributionSub, _6_VelocityDistribution.f90:60-85).            */
KAOS_HD void kaos_seed_ion(double q, double p, double sigma, double cphi, double sphi, const double u[6],
                           double* rs, double* vpar, double* vperp1, double* vperp2, double* vperp) {
  double r = kaos_rs_of_q(q, p);
  kaos_line_t L = kaos_line(r, p);
  *rs = r;
  kaos_mb_project(u, sigma, cphi, sphi, L, vpar, vperp1, vperp2, vperp);
}

/* ------------------------------------------------------------------------- */
/* ICR wave heating (WaveHeatingSub, _8_1_1_0_WaveHeating.f90:32-314).       */

/* exp(x) for |x| < 700: x = k ln2 + r, |r| <= ln2/2 (ln2 in two parts, the
 * same split as kaos_log), exp(r) by its Taylor series to r^14/14!, then
 * the exact scaling by 2^k.                                                 */
KAOS_HD double kaos_exp(double x) {
  double k = rint(x * 1.44269504088896338700);
  double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
  double p = 1.0 / 87178291200.0;                 /* 1/14! */
  p = p * r + 1.0 / 6227020800.0; p = p * r + 1.0 / 479001600.0; p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;    p = p * r + 1.0 / 362880.0;    p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;       p = p * r + 1.0 / 720.0;       p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;         p = p * r + 1.0 / 6.0;         p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
KAOS_HD void kaos_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The uniforms in (0, 1] of one ion's two Gaussian kicks at step n
 * (GaussianRNGSub twice in the reference, _8_1_1_0_WaveHeating.f90:56-64;
 * here one Box-Muller pair from u[0], u[1]): Philox4x32-10 keyed by the seed,
 * counter (low 32 bits of the ion id, step, tube, "ICR" xor the id's high 32
 * bits: ids below 2^32 leave the last word "ICR0").                         */
KAOS_HD void kaos_icr_uniforms(uint64_t seed, int64_t step, int tube, int64_t id, double u[4]) {
  uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) ^ 0x49435248u };
  uint32_t ctr[4] = { (uint32_t)id, (uint32_t)step, (uint32_t)tube, 0x49435230u ^ (uint32_t)((uint64_t)id >> 32) };
  uint32_t o[4];
  kaos_philox4x32(ctr, key, o);
  for (int i = 0; i < 4; ++i) u[i] = ((double)o[i] + 1.0) * (1.0 / 4294967296.0);
}

/* Heating parameters of a tube (the per-cell arrays of SimParameterizationSub,
 * _1_SimParameterization.f90:1430-1450, are uniform over the cells). */
typedef struct {
  double qs, ms, h;
  double lambda_perp;    /* lambdaPerpp: 0 = infinite-wavelength limit   */
  double eta_lh;         /* EtaLHp                                       */
  double xi1, xi2;       /* XiPerp1p, XiPerp2p                           */
  double s0;             /* S0p [(V^2/m^2)/Hz]                           */
  double omega_g0;       /* 2 pi f0p [rad/s]                             */
  double chi1, chi2;     /* ChiPerp1p, ChiPerp2p                         */
  double sum_ds;         /* sum of dsICRGT = hqC dqC over the cells      */
} kaos_icr_t;

/* x^(-chi) for x > 0: exactly 1/(x ... x) for a whole chi up to 8, else
 * exp(-chi ln x). */
KAOS_HD double kaos_pow_neg(double x, double chi) {
  if (chi == rint(chi) && fabs(chi) <= 8.0) {
    int k = (int)fabs(chi);
    double p = 1.0;
    for (int i = 0; i < k; ++i) p = p * x;
    return (chi >= 0.0) ? 1.0 / p : p;
  }
  return kaos_exp(-chi * kaos_log(x));
}

/* Stochastic ICR kick of one ion (ICRCOHERENCEflag = 0) at the stage-A point
 * of line geometry L (r, w) with input parallel velocity vpar and
 * perpendicular components (v1, v2); u: kaos_icr_uniforms.  WaveHeating:
 * SigmaPerp (:75-84), DPerp1/2 (:87-92), epsPerp and tauPerp (:100-105),
 * the kicks (:140-151), Vperp1N/2N and VperpN (:296-303).  Gyrofrequency
 * OmegaG = qs B / ms with B = m sqrt(ell) / r^3 (_KineticMainParams.f90:621-633),
 * ell = 1 + 3 cos^2 theta = 4 - 3 w. */
KAOS_HD void kaos_icr_kick(const kaos_icr_t* P, double r, double w, double vpar, double v1, double v2,
                           const double u[4], double* v1N, double* v2N, double* vpN) {
  const double R = sqrt(-2.0 * kaos_log(u[0]));
  const double g1 = R * kaos_cos2pi(u[1]), g2 = R * kaos_sin2pi(u[1]);
  const double B = (KAOS_MDIP * sqrt(4.0 - 3.0 * w)) / ((r * r) * r);
  const double og = (P->qs * B) / P->ms;
  const double vperp = sqrt(v1 * v1 + v2 * v2);
  double sig = 1.0;
  if (P->lambda_perp != 0.0 && ((KAOS_TWOPI * vperp) / og) >= P->lambda_perp)
    sig = (KAOS_TWOPI * vperp) / (P->lambda_perp * og);
  const double isig3 = 1.0 / ((sig * sig) * sig);
  const double x = og / P->omega_g0;
  const double c = (P->qs * P->qs) / (2.0 * (P->ms * P->ms));
  const double D1 = (((c * (P->xi1 * P->eta_lh)) * P->s0) * kaos_pow_neg(x, P->chi1)) * isig3;
  const double D2 = (((c * (P->xi2 * P->eta_lh)) * P->s0) * kaos_pow_neg(x, P->chi2)) * isig3;
  const double eps = fabs(fabs(vpar) / (og * P->sum_ds));
  const double tau = 1.0 / ((og / KAOS_TWOPI) * fabs(sqrt(eps)));
  double dv1 = 0.0, dv2 = 0.0;
  if (P->h <= tau) {
    dv1 = ((P->h / tau) * sqrt((2.0 * D1) * tau)) * g1;
    dv2 = ((P->h / tau) * sqrt((2.0 * D2) * tau)) * g2;
  }
  *v1N = v1 + dv1;
  *v2N = v2 + dv2;
  *vpN = fabs(sqrt((*v1N) * (*v1N) + (*v2N) * (*v2N)));
}

/* ------------------------------------------------------------------------- */
/* Per-tube field arithmetic (run once per master step per tube).            */

/* MomentFilterSub, _8_7_1_MomentFilter.f90:25-147: asymmetric sliding average
 * with window W (odd) and shrinking one-sided windows near LB/UB.  in/out are
 * 0-based arrays of the nq cells.                                          */
KAOS_HD double kaos_moment_filter_cell(const double* in, int nq, int W, int Q) {
  int hw = (W - 1) / 2;
  int LB = 1, UB = nq;
  double s1, s2, a1, a2, out = 0.0;
  if (Q <= LB - 1 + hw) {
    s1 = in[Q - 1] + in[Q];
    for (int k = 2; k <= hw; ++k) s1 = s1 + in[Q + k - 1];
    a1 = s1 / (double)(hw + 1);
    if (Q == LB) out = a1;
    else if (Q == LB + 1) { a2 = in[LB - 1]; out = (a1 + a2) / 2.0; }
    else {
      s2 = in[Q - 2];
      for (int k = 2; k <= Q - LB; ++k) s2 = s2 + in[Q - k - 1];
      a2 = s2 / (double)(Q - LB);
      out = (a1 + a2) / 2.0;
    }
  }
  if (Q > LB - 1 + hw && Q < UB - hw) {
    s1 = in[Q - 1] + in[Q];
    for (int k = 2; k <= hw; ++k) s1 = s1 + in[Q + k - 1];
    s2 = in[Q - 2];
    for (int k = 2; k <= hw; ++k) s2 = s2 + in[Q - k - 1];
    a1 = s1 / (double)(hw + 1);
    a2 = s2 / (double)hw;
    out = (a1 + a2) / 2.0;
  }
  if (Q >= UB - hw) {
    s2 = in[Q - 1] + in[Q - 2];
    for (int k = 2; k <= hw; ++k) s2 = s2 + in[Q - k - 1];
    a2 = s2 / (double)(hw + 1);
    if (Q == UB) out = a2;
    else if (Q == UB - 1) { a1 = in[UB - 1]; out = (a1 + a2) / 2.0; }
    else {
      s1 = in[Q];
      for (int k = 2; k <= UB - Q; ++k) s1 = s1 + in[Q + k - 1];
      a1 = s1 / (double)(UB - Q);
      out = (a1 + a2) / 2.0;
    }
  }
  return out;
}

KAOS_HD void kaos_moment_filter(const double* in, double* out, int nq, int W) {
  for (int Q = 1; Q <= nq; ++Q) out[Q - 1] = kaos_moment_filter_cell(in, nq, W, Q);
}

/* AmbipolarEfieldSub electron-pressure field with the moment filter on,
 * self-consistent density and no inertial term (_8_7_AmbipolarEfield.f90:
 * 34-82 filter of M0(nn-1); 622-1269 stencils: 5-point one-sided at LB/UB,
 * 3-point central at LB+1/UB-1, 5-point central inside; 1276-1287 LB..LB+2
 * copy of LB+3; 1297-1299 EAmag = EAInertial(=0) + EAPressure).  south != 0
 * when qGL(1) <= 0.  M0F receives the filtered density M0FiltAvrgRT.       */
/* Stencil part: EA from the filtered density n (M0FiltAvrgRT) and the raw
 * previous-step density M0prev (which only gates cells with no ions). */
KAOS_HD double kaos_ea_cell(int Q, int nq, int south, int eamb, int eambsign, double qs,
                            const double* hqC, const double* dqC, const double* Te,
                            const double* M0prev, const double* n) {
  if (M0prev[Q] == 0.0 || !eamb) return 0.0;
  double dq = dqC[Q];
  double R1 = (1.0 / fabs(hqC[Q])) * ((KAOS_KB * Te[Q]) / (qs * n[Q]));
  double R2;
  if (Q == 0) {
    double S = (((-25.0 * n[Q] + 48.0 * n[Q + 1]) - 36.0 * n[Q + 2]) + 16.0 * n[Q + 3]) - 3.0 * n[Q + 4];
    if (!eambsign) R2 = fabs(S) / (12.0 * dq);
    else R2 = south ? (-S) / (-12.0 * dq) : (-S) / (12.0 * dq);
  } else if (Q == nq - 1) {
    double S = (((-25.0 * n[Q] + 48.0 * n[Q - 1]) - 36.0 * n[Q - 2]) + 16.0 * n[Q - 3]) - 3.0 * n[Q - 4];
    if (!eambsign) R2 = fabs(S) / (-12.0 * dq);
    else R2 = south ? (-S) / (12.0 * dq) : (-S) / (-12.0 * dq);
  } else if (Q == 1 || Q == nq - 2) {
    double S = n[Q + 1] - n[Q - 1];
    if (!eambsign) R2 = fabs(S) / (2.0 * dq);
    else R2 = south ? (-S) / (-2.0 * dq) : (-S) / (2.0 * dq);
  } else {
    double S = ((n[Q - 2] - 8.0 * n[Q - 1]) + 8.0 * n[Q + 1]) - n[Q + 2];
    if (!eambsign) R2 = fabs(S) / (12.0 * dq);
    else R2 = south ? (-S) / (-12.0 * dq) : (-S) / (12.0 * dq);
  }
  double P = R1 * R2;
  return eambsign ? P : fabs(P);
}

/* LB, LB+1, LB+2 take the LB+3 value where they hold ions (:1276-1287). */
KAOS_HD void kaos_ea_override(double* EA, const double* M0prev, int eamb) {
  double e3 = EA[3];
  for (int Q = 0; Q < 3; ++Q)
    if (M0prev[Q] != 0.0 && eamb) EA[Q] = e3;
}

/* Stencil part: EA from the filtered density n (M0FiltAvrgRT) and the raw
 * previous-step density M0prev (which only gates cells with no ions). */
KAOS_HD void kaos_ambipolar_stencil(int nq, int south, int eamb, int eambsign, double qs,
                                    const double* hqC, const double* dqC, const double* Te,
                                    const double* M0prev, const double* n, double* EA) {
  for (int Q = 0; Q < nq; ++Q)
    EA[Q] = kaos_ea_cell(Q, nq, south, eamb, eambsign, qs, hqC, dqC, Te, M0prev, n);
  kaos_ea_override(EA, M0prev, eamb);
}

KAOS_HD void kaos_ambipolar(int nq, int south, int eamb, int eambsign, int W, double qs,
                            const double* hqC, const double* dqC, const double* Te,
                            const double* M0prev, double* M0F, double* EA) {
  kaos_moment_filter(M0prev, M0F, nq, W);
  kaos_ambipolar_stencil(nq, south, eamb, eambsign, qs, hqC, dqC, Te, M0prev, M0F, EA);
}

/* GravfieldSub cell value, _8_7_2_Gravfield.f90:57-60. */
KAOS_HD double kaos_grav_cell(int grav, double ms, double qs, double cth, double rc, double ellc) {
  if (!grav) return 0.0;
  return (-2.0 * KAOS_GG * KAOS_ME * ms * cth) / (qs * (rc * rc) * sqrt(ellc));
}

/* Particle field values from a cell-interpolated E (_8_7_AmbipolarEfield.f90:
 * 1477-1493 with EAMBSIGN, _8_7_2_Gravfield.f90:156-162). */
KAOS_HD double kaos_aea_of(double e_interp, double qsms, int eambsign) {
  double e = eambsign ? e_interp : fabs(e_interp);
  if (e == 0.0) return 0.0;
  return eambsign ? qsms * e : fabs(qsms * e);
}
KAOS_HD double kaos_ag_of(double e_interp, double qsms) {
  double e = fabs(e_interp);
  if (e == 0.0) return 0.0;
  return fabs(qsms * e);
}

/* ------------------------------------------------------------------------- */
/* Moments of one config cell from its 2-D phase-space counts (row-major
 * [Vperpind-1][Vparind-1]): IonParticleCountsSub normalisation
 * (_8_3_1:118-121, 713-725), IonDistributionFunctionsSub F = NphRT/d3v
 * (_8_4_1:148-158), NonUniform2DIntegratorSub MM = 2 pi sum dV g
 * (_8_5:29-43) and the integrands of _8_6_1.._8_6_6 (2-D branch, centring
 * flags perp = 0, par = 1).
 *
 * One pass over the counts: within a Vpar column b, V||c and dVpar are
 * constant, so every integrand is a column aggregate times a column factor:
 *   A0 = sum_a dV vc F,  A1 = sum_a dV vc^2 F,  A3 = sum_a dV vc^3 F
 * (dV = dVperp_a dVpar_b, vc = Vperp centre), and
 *   M0 ~ sum_b A0, M1perp ~ sum_b A1, M1par ~ sum_b uc A0, M2perp ~ sum_b A3,
 *   M2 ~ sum_b (A3 + d^2 A0), M2par ~ sum_b d^2 A0,  d = uc - M1par,
 * with the reference's fallback to the uncentred integrand where the centred
 * one is zero or NaN (_8_6_4:292-345, _8_6_6:272-299) applied per column.
 * Summation order (fixed): rows a = g, g + G, g + 2G, ... for row group
 * g = 0..G-1 (G = kaos_mom_groups), the G partials added in g order, then the
 * columns by the pairwise tree kaos_tree_sum.                               */

#define KAOS_MOM_LANES 256

typedef struct {
  int nvpar, nvperp;
  const double *d3v, *dvperp, *dvpar, *vperpGC, *vparGC;
} kaos_vgrid_t;

/* Row groups G: one (each column summed over all rows in order) when a
 * wavefront holds every column (nvpar <= 64), else KAOS_MOM_LANES / nvpar. */
KAOS_HD int kaos_mom_groups(int nvpar) {
  if (nvpar <= 64) return 1;
  int g = KAOS_MOM_LANES / nvpar;
  return g < 1 ? 1 : g;
}

/* Partial column aggregates of column b over rows g0, g0 + G, ... within
 * [a0, a1], and the partial count of those cells (empty cells contribute
 * nothing, so rows left out that hold no counts change nothing).  The device
 * compile unrolls the row loop (4 rows) so their loads issue together; the
 * arithmetic order is the same on both compilers. */
#if defined(__HIP_DEVICE_COMPILE__)
#ifndef KAOS_MOM_UNROLL
#define KAOS_MOM_UNROLL 4
#endif
#define KAOS_PRAGMA_(x) _Pragma(#x)
#define KAOS_UNROLL_N_(n) KAOS_PRAGMA_(unroll n)
#define KAOS_MOM_UNROLL_PRAGMA KAOS_UNROLL_N_(KAOS_MOM_UNROLL)
#else
#define KAOS_MOM_UNROLL_PRAGMA
#endif
KAOS_HD void kaos_mom_part(const kaos_vgrid_t* v, const uint32_t* cnt, int b, int g0, int G, int a0, int a1,
                           double k, double P[3], uint32_t* csum) {
  double s0 = 0.0, s1 = 0.0, s3 = 0.0;
  uint32_t cs = 0;
  const double dvb = v->dvpar[b];
  int a = a0 + (((g0 - a0) % G) + G) % G;   /* first row >= a0 of this group */
  KAOS_MOM_UNROLL_PRAGMA
  for (; a <= a1; a += G) {
    int i = a * v->nvpar + b;
    /* every load of the row first (all in range), so that the unrolled rows'
     * loads are in flight together; the arithmetic is unchanged */
    const uint32_t c = cnt[i];
    const double d3 = v->d3v[i], dvp = v->dvperp[a], vc = v->vperpGC[a];
    cs += c;
    if (c != 0u) {
      double f = ((double)c * k) / d3;
      double w = dvp * dvb;
      s0 = s0 + w * (vc * f);
      s1 = s1 + w * ((vc * vc) * f);
      s3 = s3 + w * ((vc * (vc * vc)) * f);
    }
  }
  P[0] = s0; P[1] = s1; P[2] = s3;
  *csum = cs;
}

/* Pairwise tree sum of x[0..n) in place (x[0] holds the result): at stride
 * s = 1, 2, 4, ... x[i] += x[i + s] for i = 0, 2s, 4s, ...                  */
KAOS_HD void kaos_tree_sum(double* x, int n) {
  for (int st = 1; st < n; st *= 2)
    for (int i = 0; i + st < n; i += 2 * st) x[i] = x[i] + x[i + st];
}

/* Column values of the first stage: [A0, A1, uc A0, A3]. */
KAOS_HD void kaos_mom_col_first(double uc, const double A[3], double o[4]) {
  o[0] = A[0]; o[1] = A[1]; o[2] = uc * A[0]; o[3] = A[2];
}

/* M0, M1perp, M1par, M2perp from the first-stage sums S[0..3]. */
KAOS_HD void kaos_mom_finish1(const double S[4], double ms, double M[6]) {
  double M0 = (2.0 * KAOS_PI) * S[0];
  M[0] = M0;
  M[1] = (M0 == 0.0) ? 0.0 : ((2.0 * KAOS_PI) * S[1]) / M0;
  M[2] = (M0 == 0.0) ? 0.0 : ((2.0 * KAOS_PI) * S[2]) / M0;
  M[4] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (((2.0 * KAOS_PI) * S[3]) / M0);
}

/* Column values of the second stage (centred by M1par): [M2, M2par] integrands. */
KAOS_HD void kaos_mom_col_second(double uc, const double A[3], double M1par, double o[2]) {
  double d = uc - M1par;
  double c3 = A[2] + (d * d) * A[0];
  if (c3 != c3 || c3 == 0.0) c3 = A[2] + (uc * uc) * A[0];
  double c5 = (d * d) * A[0];
  if (c5 != c5 || c5 == 0.0) c5 = (uc * uc) * A[0];
  o[0] = c3; o[1] = c5;
}

/* M2, M2par from the second-stage sums. */
KAOS_HD void kaos_mom_finish2(const double S[2], double ms, double M[6]) {
  double M0 = M[0];
  M[3] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (((2.0 * KAOS_PI) * S[0]) / M0);
  M[5] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (((2.0 * KAOS_PI) * S[1]) / M0);
}

/* ------------------------------------------------------------------------- */
/* The 3-D velocity grid (ION2VPERPflag = 1): Vperp1 x Vperp2 x Vpar cells.   */
/* Counts N2PerpphReNormT(nn, V1, V2, Vpar) (_8_3_1_IonParticleCounts.f90:    */
/* 336-361), F = N2PerpphRT / d3v with d3v = dV1 dV2 dVpar                    */
/* (_8_4_1_IonDistributionFunctions.f90:50-60, _0_0_VelGridGenerator.f90:     */
/* 395-398), and the moments of NonUniform3DIonIntegratorSub                  */
/* (_8_5_1_NonUniform3DIonIntegrator.f90:25-48: sum of ((dV1 dV2) dVpar) g,   */
/* no 2 pi):                                                                  */
/*   M0 = sum F;  M1Perp1 = sum V1c F / M0;  M1Perp2 = sum V2c F / M0;        */
/*   M1Par = sum Vparc F / M0 (_8_6_1..3);                                    */
/*   M2Perp1 = (m/2) sum V1c^2 F / M0, M2Perp2 likewise (uncentred,           */
/*   PHASEENERGYPERPIONMOMENTCENTERflag = 0, _8_6_5);                         */
/*   M2 = (m/2) sum (V1c^2 + V2c^2 + (Vparc - M1Par)^2) F / M0 (_8_6_4:82-120) */
/*   and M2Par = (m/2) sum (Vparc - M1Par)^2 F / M0 (_8_6_6:82-95), each      */
/*   falling back to the uncentred integrand where the centred one is 0 or   */
/*   NaN (applied per Vpar column, as the 2-D moments do).                    */
/* Counts are [Q][V1][V2][Vpar], Vpar fastest; a row is one (V1, V2) pair,    */
/* r = i1 nv2 + i2.  Summation order (fixed): column b over rows g, g + G,   */
/* ... (G = kaos_mom3_groups), the G partials added in g order, then the      */
/* columns by kaos_tree_sum.                                                  */

typedef struct {
  int nvpar, nv1, nv2;
  const double *d3v, *dv1, *dv2, *dvpar, *v1GC, *v2GC, *vparGC;
} kaos_vgrid3_t;

/* Lanes of the 3-D moment sums: one 1024-thread block per config cell, so
 * (V1, V2) rows split into 1024 / nvpar groups. */
#define KAOS_MOM3_LANES 1024
KAOS_HD int kaos_mom3_groups(int nvpar) {
  int g = KAOS_MOM3_LANES / nvpar;
  return g < 1 ? 1 : g;
}

/* Column b's partial sums over rows g0, g0 + G, ... within [r0, r1]:
 * P = {sum w f, sum w V1c f, sum w V2c f, sum w V1c^2 f, sum w V2c^2 f} with
 * w = (dV1 dV2) dVpar and f = (N k) / d3v, and the partial count.          */
KAOS_HD void kaos_mom3_part(const kaos_vgrid3_t* v, const uint32_t* cnt, int b, int g0, int G, int r0, int r1,
                            double k, double P[5], uint32_t* csum) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  uint32_t cs = 0;
  const double dvb = v->dvpar[b];
  int r = r0 + (((g0 - r0) % G) + G) % G;   /* first row >= r0 of this group */
  KAOS_MOM_UNROLL_PRAGMA
  for (; r <= r1; r += G) {
    const int i = r * v->nvpar + b;
    /* every load of the row first (all in range), so that the unrolled rows'
     * loads are in flight together; the arithmetic is unchanged */
    const uint32_t c = cnt[i];
    const int i1 = r / v->nv2, i2 = r - i1 * v->nv2;
    const double d3 = v->d3v[i], dv1 = v->dv1[i1], dv2 = v->dv2[i2];
    const double u1 = v->v1GC[i1], u2 = v->v2GC[i2];
    cs += c;
    if (c != 0u) {
      double f = ((double)c * k) / d3;
      double w = (dv1 * dv2) * dvb;
      s0 = s0 + w * f;
      s1 = s1 + w * (u1 * f);
      s2 = s2 + w * (u2 * f);
      s3 = s3 + w * ((u1 * u1) * f);
      s4 = s4 + w * ((u2 * u2) * f);
    }
  }
  P[0] = s0; P[1] = s1; P[2] = s2; P[3] = s3; P[4] = s4;
  *csum = cs;
}

/* Column values of the first stage: [A0, A1, A2, uc A0, A3, A4]. */
KAOS_HD void kaos_mom3_col_first(double uc, const double A[5], double o[6]) {
  o[0] = A[0]; o[1] = A[1]; o[2] = A[2]; o[3] = uc * A[0]; o[4] = A[3]; o[5] = A[4];
}

/* M[8] = M0, M1Perp1, M1Perp2, M1Par, M2, M2Perp1, M2Perp2, M2Par: the
 * first-stage ones from the summed column values S[6].                     */
KAOS_HD void kaos_mom3_finish1(const double S[6], double ms, double M[8]) {
  double M0 = S[0];
  M[0] = M0;
  M[1] = (M0 == 0.0) ? 0.0 : S[1] / M0;
  M[2] = (M0 == 0.0) ? 0.0 : S[2] / M0;
  M[3] = (M0 == 0.0) ? 0.0 : S[3] / M0;
  M[5] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (S[4] / M0);
  M[6] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (S[5] / M0);
}

/* Column values of the second stage (centred by M1Par): [M2, M2Par] integrands. */
KAOS_HD void kaos_mom3_col_second(double uc, const double A[5], double M1par, double o[2]) {
  double d = uc - M1par;
  double c2 = (A[3] + A[4]) + (d * d) * A[0];
  if (c2 != c2 || c2 == 0.0) c2 = (A[3] + A[4]) + (uc * uc) * A[0];
  double cp = (d * d) * A[0];
  if (cp != cp || cp == 0.0) cp = (uc * uc) * A[0];
  o[0] = c2; o[1] = cp;
}

KAOS_HD void kaos_mom3_finish2(const double S[2], double ms, double M[8]) {
  double M0 = M[0];
  M[4] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (S[0] / M0);
  M[7] = (M0 == 0.0) ? 0.0 : (ms / 2.0) * (S[1] / M0);
}

/* The three velocity cells of an ion on the 3-D grid (KUA:985-1080): cell 1
 * is [GL, GH], the others (GL, GH]; 0 when off the grid (no stale index).  */
KAOS_HD void kaos_vcells3(double vpar, double v1, double v2, const double* vparGH, double vparGL0, int nvpar,
                          const double* v1GH, double v1GL0, int nv1, const double* v2GH, double v2GL0, int nv2,
                          int* ip, int* i1, int* i2) {
  *ip = kaos_find_vcell(vpar, vparGH, vparGL0, nvpar);
  *i1 = kaos_find_vcell(v1, v1GH, v1GL0, nv1);
  *i2 = kaos_find_vcell(v2, v2GH, v2GL0, nv2);
}

/* ------------------------------------------------------------------------- */
/* Ion-neutral charge exchange (QEXCHANGEflagT = 1): collision counts per     */
/* config cell (IonNeutralCollisionFrequencyA/BSub), the ion -> ENA           */
/* conversion (IonNeutralCollisionsSub) and the ballistic ENA push under the  */
/* neutral's gravity (the ENA branches of KineticUpdateA/B/C and              */
/* KineticRK4UpdateSub).                                                      */

#define KAOS_CX_MAXITER 4096      /* bound of the Poisson product loop         */

/* exp(x) for x <= 0, 0 below the binary64 range. */
KAOS_HD double kaos_exp_neg(double x) { return (x < -745.2) ? 0.0 : kaos_exp(x); }

/* Four uniforms of a cell's collision stream: Philox4x64-10 keyed by (seed,
 * "CXNT") with counter (cell Q, master step n, tube, block).  Blocks 0-1
 * give the neutral velocity's three Box-Muller pairs, blocks 2.. the Poisson
 * loop's uniforms. */
KAOS_HD void kaos_cx_block(uint64_t seed, int64_t step, int tube, int Q, uint64_t blk, double u[4]) {
  uint64_t key[2] = { seed, 0x43584E54ULL };
  uint64_t ctr[4] = { (uint64_t)Q, (uint64_t)step, (uint64_t)tube, blk };
  uint64_t o[4];
  kaos_philox4x64(ctr, key, o);
  u[0] = kaos_u01(o[0]); u[1] = kaos_u01(o[1]); u[2] = kaos_u01(o[2]); u[3] = kaos_u01(o[3]);
}

/* Charge-exchange cross-section from the parallel energy moment M2par (J):
 * E = 6.242e18 * 1e-3 * M2par (keV), sigma = (4.07 - 0.269 ln E)^2
 * (1 - exp(-415/E))^0.8, and 0 for M2par = 0
 * (_8_6_6_1_IonNeutralCollisionFrequencyA.f90:93-114). */
KAOS_HD double kaos_cx_sigma(double m2par) {
  if (m2par == 0.0) return 0.0;
  const double E = (6.242e18 * 1e-3) * m2par;
  const double a = 4.07 - 0.269 * kaos_log(E);
  const double b = 1.0 - kaos_exp_neg(-415.0 / E);
  return (a * a) * ((b > 0.0) ? kaos_exp(0.8 * kaos_log(b)) : 0.0);
}

/* Poisson variate of mean mu by the product of uniforms
 * (_3_2_PoissonRNG.f90:30-40): uniforms are multiplied while the product
 * stays above exp(-mu); the count is their number less one.  The uniforms
 * come from blocks 2.. of the cell's stream; at most KAOS_CX_MAXITER. */
KAOS_HD double kaos_poisson(double mu, uint64_t seed, int64_t step, int tube, int Q) {
  const double L = kaos_exp_neg(-mu);
  double prod = 1.0, k = -1.0, u[4];
  int i = 4;
  uint64_t blk = 2;
  for (int it = 0; it < KAOS_CX_MAXITER && prod > L; ++it) {
    if (i == 4) { kaos_cx_block(seed, step, tube, Q, blk++, u); i = 0; }
    prod = prod * u[i++];
    k = k + 1.0;
  }
  return k;
}

/* Inputs of one cell's collision count at a master step. */
typedef struct {
  double m0, m1par, m1p1, m1p2, m2par;  /* moments (filtered, MOMENTFILTERflagT = 1) */
  double cth, ell;                      /* cos(thetaGCGT), ellGCGT of the cell centre */
  double cphi, sphi;                    /* tube longitude */
  double nsneut;                        /* nsnormCNeutGT: neutral macro-particles of the cell */
  double h, sig_n;                      /* time step; neutral thermal speed sqrt(kB TNeut / mNeut) */
  double NqRT, d3x, nsnormfac;          /* NqRTp and the cap's normalisation */
  double nqc;                           /* NqReNormT: the cell's macro-ion count */
  double sig_scale;                     /* factor on sigma (1: the expression as written) */
} kaos_cx_in_t;

typedef struct { double sigma, nu, poi; } kaos_cx_out_t;

/* IonNeutralCollisionFrequencyASub for one cell (_8_6_6_1:93-300): sigma;
 * for a populated cell a neutral velocity of three Gaussians, the ion mean
 * velocity at the cell centre in Cartesian components, their relative
 * speed, the Poisson count of mean nsneut sigma v h capped at the cell's
 * macro-ions; -2 for an empty cell or sigma = 0.  Then the one-rank share of
 * IonNeutralCollisionFrequencyBSub (_8_6_6_2:35-77): the count, at most the
 * cell's NqReNormT, 0 unless positive (poi, converted nint(poi) times per
 * step of the next master interval). */
KAOS_HD kaos_cx_out_t kaos_cx_cell(const kaos_cx_in_t* I, uint64_t seed, int64_t step, int tube, int Q) {
  kaos_cx_out_t o;
  o.sigma = I->sig_scale * kaos_cx_sigma(I->m2par);
  o.nu = -2.0;
  if (I->m0 != 0.0) {
    double u[8];
    kaos_cx_block(seed, step, tube, Q, 0, u);
    kaos_cx_block(seed, step, tube, Q, 1, u + 4);
    const double Vnx = I->sig_n * kaos_gauss(u[0], u[1]);
    const double Vny = I->sig_n * kaos_gauss(u[2], u[3]);
    const double Vnz = I->sig_n * kaos_gauss(u[4], u[5]);
    const double c = I->cth, s = sqrt(1.0 - c * c), sl = sqrt(I->ell);
    const double prho = (3.0 * I->m1par * c * s + I->m1p1 * (1.0 - 3.0 * (c * c))) / sl;
    const double Mx = I->cphi * prho - I->m1p2 * I->sphi;
    const double My = I->sphi * prho + I->m1p2 * I->cphi;
    const double Mz = I->m1par * (3.0 * (c * c) - 1.0) / sl + 3.0 * I->m1p1 * c * s / sl;
    const double dx = Mx - Vnx, dy = My - Vny, dz = Mz - Vnz;
    const double vrel = sqrt((dx * dx + dy * dy) + dz * dz);
    if (o.sigma != 0.0) {
      const double mu = ((I->nsneut * o.sigma) * vrel) * I->h;
      double nu = kaos_poisson(mu, seed, step, tube, Q);
      const double cap = I->NqRT * (I->d3x / I->nsnormfac);
      if (nu > cap) nu = cap;
      o.nu = nu;
    }
  }
  o.poi = 0.0;
  if (o.nu > 0.0 && I->nqc != 0.0) o.poi = (o.nu <= I->nqc) ? o.nu : I->nqc;
  return o;
}

/* Ions converted per step: nint(poi). */
KAOS_HD int64_t kaos_cx_take(double poi) { return (poi > 0.0) ? (int64_t)floor(poi + 0.5) : 0; }

/* The ENA made from an ion (IonNeutralCollisionsSub, _8_1_1_1:31-58): its
 * position, and the velocity (Vp, Vq, Vphi) = (Vperp1, b-hat(thetak4) . V,
 * Vperp2) composed in the dipole frame of the stage-4 point of the ion's
 * last push (cos(thetak4) = ck4, phik4 = the tube longitude); V = V|| b-hat
 * is the ion's velocity after that push.  X = (x, y, z, Vx, Vy, Vz). */
KAOS_HD void kaos_ena_from_ion(double rs, double vs, double v1, double v2, double ck4, double p, double cphi,
                               double sphi, double X[6]) {
  double x, y, z, Vx, Vy, Vz;
  kaos_cartesian(rs, vs, p, cphi, sphi, &x, &y, &z, &Vx, &Vy, &Vz);
  const double c = ck4, s = sqrt(1.0 - c * c), sl = sqrt(1.0 + 3.0 * (c * c));
  const double vq = 3.0 * c * s * (Vx * cphi + Vy * sphi) / sl + Vz * (3.0 * (c * c) - 1.0) / sl;
  const double vrho = (3.0 * vq * c * s + v1 * (1.0 - 3.0 * (c * c))) / sl;
  X[0] = x; X[1] = y; X[2] = z;
  X[3] = cphi * vrho - v2 * sphi;
  X[4] = sphi * vrho + v2 * cphi;
  X[5] = vq * (3.0 * (c * c) - 1.0) / sl + 3.0 * v1 * c * s / sl;
}

/* The neutral's gravitational acceleration -G ME / r^2 r-hat
 * (AGENAx/y/zsub, _KineticMainParams.f90:695-714), 0 with GRAVflag off. */
KAOS_HD void kaos_ena_accel(double x, double y, double z, int grav, double a[3]) {
  if (!grav) { a[0] = a[1] = a[2] = 0.0; return; }
  const double r2 = (x * x + y * y) + z * z, r = sqrt(r2);
  const double g = (KAOS_GG * KAOS_ME) / r2;
  a[0] = -g * (x / r); a[1] = -g * (y / r); a[2] = -g * (z / r);
}

/* q of an ENA at (x, y, z): RE^2 cos(theta) / r^2 (qsub), the coordinate its
 * config cell is scanned on. */
KAOS_HD double kaos_ena_q(double x, double y, double z) {
  const double r = sqrt((x * x + y * y) + z * z);
  return (KAOS_RE2 * (z / r)) / (r * r);
}

/* One RK4 step of an ENA with the reference's stages (k2 = k3,
 * _8_1_KineticRK4Update.f90:29-156): accelerations at x, at x + h/2 V and at
 * x + h V + h^2/2 a1; no re-projection, phi free. */
KAOS_HD void kaos_ena_push(double X[6], double h, int grav) {
  const double h2 = h / 2.0, hsq2 = (h * h) / 2.0, h6 = h / 6.0;
  double a1[3], a2[3], a4[3], xn[3], vn[3];
  kaos_ena_accel(X[0], X[1], X[2], grav, a1);
  kaos_ena_accel(X[0] + h2 * X[3], X[1] + h2 * X[4], X[2] + h2 * X[5], grav, a2);
  kaos_ena_accel((X[0] + h * X[3]) + hsq2 * a1[0], (X[1] + h * X[4]) + hsq2 * a1[1],
                 (X[2] + h * X[5]) + hsq2 * a1[2], grav, a4);
  for (int i = 0; i < 3; ++i) {
    const double V = X[3 + i];
    const double k2 = V + h2 * a1[i], k4 = V + h * a2[i];
    vn[i] = V + h6 * (((a1[i] + 2.0 * a2[i]) + 2.0 * a2[i]) + a4[i]);
    xn[i] = X[i] + h6 * (((V + 2.0 * k2) + 2.0 * k2) + k4);
  }
  for (int i = 0; i < 3; ++i) { X[i] = xn[i]; X[3 + i] = vn[i]; }
}

/* ENA velocity space (QEXCHANGEflagT = 1): the ENA's (|Vp|, Vq, Vphi) at its
 * position, in the local dipole frame there (KUA:1156-1160 and the Vpar of
 * KUA:978-980: theta from z/r, phi from the ENA's own (x, y)).             */
KAOS_HD void kaos_ena_vcomp(const double X[6], double* vp, double* vq, double* vphi) {
  const double rho = sqrt(X[0] * X[0] + X[1] * X[1]);
  const double r = sqrt(rho * rho + X[2] * X[2]);
  const double c = X[2] / r, s = rho / r, sl = sqrt(1.0 + 3.0 * (c * c));
  const double cp = (rho > 0.0) ? X[0] / rho : 1.0, sp = (rho > 0.0) ? X[1] / rho : 0.0;
  const double vrho = X[3] * cp + X[4] * sp;
  *vq = 3.0 * c * s * vrho / sl + X[5] * (3.0 * (c * c) - 1.0) / sl;
  *vp = fabs((1.0 - 3.0 * (c * c)) * vrho / sl + 3.0 * X[5] * c * s / sl);
  *vphi = -X[3] * sp + X[4] * cp;
}

/* The ENA grid of a tube: Vp, Vq, Vphi cells with their centres and widths,
 * and, for the config cell at hand, per bin [ip][iq][iphi] (Vphi fastest) the
 * metric product hVp hVq hVphi and d33v = hVp hVq hVphi dVp dVq dVphi
 * (_0_0_VelGridGenerator.f90:820-857; both depend on the cell's thetaGC). */
typedef struct {
  int nvp, nvq, nvphi;
  const double *dvp, *dvq, *dvphi, *vpGC, *vqGC, *vphiGC, *hh, *d33v;
} kaos_enagrid_t;

#define KAOS_ENA_LANES 256

/* Per-lane partial sums of a config cell's ENA moments over the bins b =
 * lane, lane + 256, ... in order: F = (count nsnormfac/d3x) / d33v
 * (ENADistributionFunctionsSub, _8_4_2:24-36), weight dVp dVq dVphi hh F
 * (NonUniform3DIntegratorSub, _8_9:25-35); stage 1 sums w, w Vp, w Vq,
 * w Vphi (ZerothENAMomentSub, FirstP/Q/PHIENAMomentSub), stage 2 sums w
 * |dV|^2, w dVp^2, w dVq^2, w dVphi^2 of the offsets from the first moments
 * m1 (SecondENAMomentSub, SecondP/Q/PHIENAMomentSub, _8_10_5:30-40). */
KAOS_HD void kaos_ena_mom_part(const kaos_enagrid_t* g, const uint32_t* cnt, int lane, double k, int stage,
                               const double m1[3], double o[4]) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const int nb = g->nvp * g->nvq * g->nvphi;
  for (int b = lane; b < nb; b += KAOS_ENA_LANES) {
    const uint32_t c = cnt[b];
    if (c == 0u) continue;
    const int iphi = b % g->nvphi, iq = (b / g->nvphi) % g->nvq, ip = b / (g->nvphi * g->nvq);
    const double F = ((double)c * k) / g->d33v[b];
    const double w = ((g->dvp[ip] * g->dvq[iq]) * g->dvphi[iphi]) * (g->hh[b] * F);
    if (stage == 1) {
      s0 = s0 + w;
      s1 = s1 + w * g->vpGC[ip];
      s2 = s2 + w * g->vqGC[iq];
      s3 = s3 + w * g->vphiGC[iphi];
    } else {
      const double dp = m1[0] - g->vpGC[ip], dq = m1[1] - g->vqGC[iq], df = m1[2] - g->vphiGC[iphi];
      s0 = s0 + w * ((dp * dp + dq * dq) + df * df);
      s1 = s1 + w * (dp * dp);
      s2 = s2 + w * (dq * dq);
      s3 = s3 + w * (df * df);
    }
  }
  o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
}

/* M0phENART, M1PphENART, M1QphENART, M1PHIphENART from the stage-1 sums
 * (0 without ENAs in the cell). */
KAOS_HD void kaos_ena_mom_finish1(const double S[4], double M[8]) {
  const double M0 = S[0];
  M[0] = M0;
  M[1] = (M0 == 0.0) ? 0.0 : S[1] / M0;
  M[2] = (M0 == 0.0) ? 0.0 : S[2] / M0;
  M[3] = (M0 == 0.0) ? 0.0 : S[3] / M0;
}

/* M2phENART, M2PphENART, M2QphENART, M2PHIphENART [eV]: 6.242e18 (ms/2) S / M0. */
KAOS_HD void kaos_ena_mom_finish2(const double S[4], double ms, double M[8]) {
  const double M0 = M[0];
  for (int i = 0; i < 4; ++i) M[4 + i] = (M0 == 0.0) ? 0.0 : (6.242e18 * (ms / 2.0)) * (S[i] / M0);
}

/* ------------------------------------------------------------------------- */
/* Grid set-up of one tube from the model inputs: ConfigGridGeneratorSub     */
/* (_0_ConfigGridGenerator.f90), ConfigGridGeneratorSubB                     */
/* (_0_1_1_ConfigGridGeneratorB.f90), VelGridGeneratorSub                    */
/* (_0_0_VelGridGenerator.f90), with NqLB = 1 and every grid-point stride    */
/* (dq, ddVpar, ...) 1, the only values for which the reference's indexing   */
/* stays inside its arrays.                                                  */

/* sin(x) for |x| < 2^20 pi/2: Cody-Waite reduction by pi/2 with a three-part
 * constant (the split of fdlibm's __ieee754_rem_pio2: 33 + 33 + 53 bits, so
 * k P1 and k P2 are exact), then the Taylor polynomials on [-pi/4, pi/4].   */
KAOS_HD double kaos_sin(double x) {
  const double k = rint(x * 6.36619772367581382433e-01);           /* 2/pi */
  const double r = ((x - k * 1.57079632673412561417e+00) - k * 6.07710050630396597660e-11) -
                   k * 2.02226624879595063154e-21;
  const int q = (int)((long long)k & 3);
  if (q == 0) return kaos_sin_poly(r);
  if (q == 1) return kaos_cos_poly(r);
  if (q == 2) return -kaos_sin_poly(r);
  return -kaos_cos_poly(r);
}

/* exp over the whole binary64 range (0 below -745.2, +inf above 709.78). */
KAOS_HD double kaos_exp_full(double x) {
  if (x < -745.2) return 0.0;
  if (x > 709.78) return HUGE_VAL;
  return kaos_exp(x);
}

/* q of grid point i (1..ngp, ngp = NqUB - NqLB + 4): uniform steps of
 * |qGB - qGA| / ngp from qGA, toward qGB, and qGB itself at the last point
 * (_0_ConfigGridGenerator.f90:299-310).                                     */
KAOS_HD double kaos_gg_qpoint(double qGA, double qGB, long long ngp, long long i) {
  if (i == ngp) return qGB;
  const double step = fabs(qGB - qGA) / (double)ngp;
  return (qGA <= 0.0) ? qGA + ((double)i - 1.0) * step : qGA - ((double)i - 1.0) * step;
}

/* One cell of the set-up arrays ("0" arrays, Qind = 1..NqUB + 2, cell 1 the
 * lower ghost; the loop's cell Q is set-up cell Q + 1,
 * _1_SimParameterization.f90:1156-1175).                                    */
typedef struct {
  double qGL, qGH, qGC, r, cth, sth, ell, hq, hp, hphi, dq, d3x;
} kaos_gg_cell_t;

/* Edges at points Qind and Qind + 1, centre, the grid root solve at the
 * centre (GridDipolePolynomialSolverSub: r of the quartic, sin(theta) =
 * sqrt(r/(RE p)), theta -> pi - theta for q <= 0), ellGC0, the metric
 * factors hqC0, hpC0, hphiC0, dqC0 and d3xC0 = hq hp hphi dq dp dphi with
 * dp = 1e-3, dphi = pi/180 (1e-9 if 0) (_0_ConfigGridGenerator.f90:433-547). */
KAOS_HD void kaos_gg_cell(double qGA, double qGB, long long ngp, double L, long long Qind, kaos_gg_cell_t* c) {
  c->qGL = kaos_gg_qpoint(qGA, qGB, ngp, Qind);
  c->qGH = kaos_gg_qpoint(qGA, qGB, ngp, Qind + 1);
  c->qGC = (c->qGL + c->qGH) / 2.0;
  c->r = kaos_root_r(c->qGC, L);
  const double w = c->r / (KAOS_RE * L);
  c->sth = sqrt(w);
  c->cth = (c->qGC <= 0.0) ? -sqrt(1.0 - w) : sqrt(1.0 - w);
  c->ell = 1.0 + 3.0 * (c->cth * c->cth);
  const double se = sqrt(c->ell);
  c->hq = fabs(((c->r * c->r) * c->r) / (KAOS_RE2 * se));
  c->hp = fabs((KAOS_RE * ((c->sth * c->sth) * c->sth)) / se);
  c->hphi = fabs(c->r * c->sth);
  c->dq = fabs(c->qGH - c->qGL);
  c->d3x = ((((c->hq * c->hp) * c->hphi) * c->dq) * 1e-3) * (KAOS_PI / 180.0);
  if (c->d3x == 0.0) c->d3x = 1e-9;
}

/* The cell's term ICbbp of the field-aligned gravitational potential
 * (_0_1_1_ConfigGridGeneratorB.f90:46-48). */
KAOS_HD double kaos_gg_icbbp(const kaos_gg_cell_t* c) {
  return ((2.0 * KAOS_GG) * KAOS_ME) * (((c->hq * c->dq) * c->cth) / ((c->r * c->r) * sqrt(c->ell)));
}

/* Hydrostatic ion density nsC = ns0 exp(-+ms gC / (kB (TsPar + Te))) of a
 * cell whose potential sum is gC = ICbb - IC0bb; sign + when qGA <= 0
 * (_0_1_1_ConfigGridGeneratorB.f90:50-73). */
KAOS_HD double kaos_gg_ns(double gC, double ms, double tpar, double te, double ns0, double qGA) {
  const double a = (ms * gC) / (KAOS_KB * (tpar + te));
  return ns0 * kaos_exp_full((qGA <= 0.0) ? a : -a);
}

/* Neutral density nsCNeut = ns0Neut exp(-+mNeut gCNeut (1/zns0Neut - 1/r)),
 * gCNeut = G ME/(kB TNeut), sign + when qGA <= 0
 * (_0_1_1_ConfigGridGeneratorB.f90:97-115). */
KAOS_HD double kaos_gg_ns_neut(double r, double mneut, double tneut, double z0, double n0, double qGA) {
  const double g = (KAOS_GG * KAOS_ME) / (KAOS_KB * tneut);
  const double a = (mneut * g) * ((1.0 / z0) - (1.0 / r));
  return n0 * kaos_exp_full((qGA <= 0.0) ? a : -a);
}

/* Point i (1 .. 2N - 1) of a symmetric velocity grid of N = NVGpF/2 points
 * per side (SYMVPARflag = 1 V||, Vperp, Vperp1/2; _0_0_VelGridGenerator.f90:
 * 70-95, 126-151, 173-198): -GB, then steps of GB/N up to the N-th point 0,
 * mirrored: P(N + k) = -P(N - k).                                           */
KAOS_HD double kaos_gg_vsym(double N, double GB, long long i) {
  const long long n = (long long)N;
  const long long j = (i > n) ? 2 * n - i : i;
  double v;
  if (j == n) v = 0.0;
  else if (j == 1) v = -GB;
  else v = -GB + ((double)j - 1.0) * (fabs(GB - 0.0) / N);
  return (i > n) ? -v : v;
}

/* Point i (1 .. N - 1) of the one-sided V|| grid (SYMVPARflag = 0, N =
 * NVparGpF points from GA = -4 sigma to GB; _0_0_VelGridGenerator.f90:
 * 203-245), reversed and negated in the southern hemisphere (qGA <= 0).     */
KAOS_HD double kaos_gg_vlin_p1(double N, double GA, double GB, long long i) {
  return (i == (long long)N) ? GB : GA + ((double)i - 1.0) * (fabs(GB - GA) / N);
}
KAOS_HD double kaos_gg_vlin(double N, double GA, double GB, int south, long long i) {
  return south ? -kaos_gg_vlin_p1(N, GA, GB, (long long)N - i + 1) : kaos_gg_vlin_p1(N, GA, GB, i);
}

/* Point i (1 .. 2N - 1) of the ENA Vp / Vphi grid (and Vq with SYMVPARflag
 * = 1; _0_0_VelGridGenerator.f90:500-615, 636-668): P1 = 0, GB/N, 2 GB/N,
 * ... with P1(N) = GB; the first N points are -P1 reversed (-GB .. -0), the
 * rest P1(i - N) (the reference reads P1(i) there, past the end of its N
 * elements).                                                                */
KAOS_HD double kaos_gg_vena(double N, double GB, long long i) {
  const long long n = (long long)N;
  if (i <= n) return -kaos_gg_vlin_p1(N, 0.0, GB, n - i + 1);
  return kaos_gg_vlin_p1(N, 0.0, GB, i - n);
}

/* ENA metric of one bin of a config cell whose centre has (cos, sin) of
 * theta and ell (_0_0_VelGridGenerator.f90:790-857): Vr, Vtheta of the bin
 * centre, hVp = |RE sin^3(Vtheta)| / sqrt(Vell), hVq = |Vr^3| / (RE^2
 * sqrt(Vell)), hVphi = |Vr sin(Vtheta)|; hh = hVp hVq hVphi, d33v = hh dVp
 * dVq dVphi (1e-9 if 0).  Vtheta is a speed that the reference passes to
 * sin as an angle; kept as written.                                         */
KAOS_HD void kaos_gg_ena_metric(double cth, double sth, double ell, double vpC, double vqC, double dvp, double dvq,
                                double dvphi, double* hh, double* d33v) {
  const double se = sqrt(ell);
  const double vr = ((2.0 * vqC) * cth) / se + (vpC * sth) / se;
  const double vt = (vqC * sth) / se - ((2.0 * vpC) * cth) / se;
  const double vell = 1.0 + 3.0 * (cth * cth);
  const double sv = kaos_sin(vt);
  const double sve = sqrt(vell);
  const double hvp = fabs((KAOS_RE * ((sv * sv) * sv)) / sve);
  const double hvq = fabs(((vr * vr) * vr) / (KAOS_RE2 * sve));
  const double hvphi = fabs(vr * sv);
  *hh = (hvp * hvq) * hvphi;
  double d = ((((hvp * hvq) * hvphi) * dvp) * dvq) * dvphi;
  if (d == 0.0) d = 1e-9;
  *d33v = d;
}

#endif /* KAOS_DIPOLE_H */
