// kaos_mixed.h -- mixed-precision form of the push + bin kernel (included by
// kaos_hip.hip after DevTube and StepArgs).
//
// State per ion, one 32-bit word each:
//   rdi  radius: bits 16..47 of a 48-bit fixed-point offset from rref (binary64
//        radius of the centre of the cell-word cell), in units of 2^-(k+16)
//        metres, k per cell so that two cell widths fit in 47 bits
//        (resolution <= width / 2^46); cell-word flag KAOS_RFLAG: the word
//        holds the binary32 offset instead (ions far outside the domain).
//   vpi  V||: bits 16..47 of a 48-bit fixed-point offset from the centre of
//        the velocity cell in the cell word (resolution <= dVpar / 2^46);
//        flag KAOS_VFLAG (ions off the velocity grid): the word holds the
//        binary32 offset of V|| from the nearer grid edge (positive above the
//        top edge, negative below GL0).
//   lo   bits 0..15 of both offsets: radius | V|| << 16.  With 32-bit offsets
//        alone the radius is rounded by ~width/2^31 every step, which walks an
//        ion ~1e-4 m from its binary64 trajectory in 200 steps (enough to bin
//        a few ions in 1e8 in a neighbouring cell); the 48-bit offsets leave
//        the binary32 stage arithmetic as the only rounding.
//   qkf  stage-4 q of the last push before a master step (binary32; it only
//        sets the field interpolation weight, so its 1e-7 relative rounding
//        moves the interpolated field by ~1e-7 of the cell-to-cell change).
//   aef/agf AEAmag/AGmag (binary32); vrf Vperp (binary32, read for output only).
//   cell (qc+1) << 16 | Vperpind << 8 | Vparind, plus the two flags; Vparind is
//        the cell of the stored V|| (the binning index of the next push, kept
//        from the previous push when V|| is off the grid); KAOS_DEAD = dead slot.
//
// Stage arithmetic (b-hat, forces, RK4 combinations) runs in binary32.
// binary64 is kept where binary32 would lose the answer: the radius and V||
// themselves (decoded offsets, the new values), the cell search (|r| against
// the binary64 radii of the q edges, so no binary32 q meets an edge), the
// root solve for the new point
// (_8_1_4_RK4DipolePolynomialSolver.f90), whose input q^2 = A/B is assembled
// as A = RE^4 [r^2 (1 - w) + Tz], B = [r^2 + Tu]^3 from the binary64
// displacement (dR, dZ) = h v b-hat_A + h^2 (2 a1 + a2) / 6 (the first term
// binary64 with sin, cos and 1/sqrt(1 + 3 cos^2) from binary32 seeds plus one
// binary64 Newton step, ~1e-14; the h^2 force term binary32); and, in the
// other kernels, the moments and the ambipolar gradient.  The new V|| is
// b-hat_N . (v b-hat_A + dV) = v (1 - |b-hat_N - b-hat_A|^2 / 2) + b-hat_N . dV
// for unit vectors, so binary32 rounding of b-hat touches only the small
// increment dV and the tiny rotation term.

#define KAOS_DEAD 0xFFFFFFFFu

// The mixed form is not bit-exact by design, so its kernel body lets the
// compiler fuse multiply-adds (the FP64 form and the shared header stay
// uncontracted for bit-exactness against the oracle).
#ifndef KAOS_MIXED_NO_CONTRACT
#define KAOS_MIXED_CONTRACT _Pragma("clang fp contract(fast)")
#else
#define KAOS_MIXED_CONTRACT
#endif
#define KAOS_RFLAG (1u << 30)
#define KAOS_VFLAG (1u << 29)
#define KAOS_QMASK 0x3FFu

__device__ __forceinline__ float kf_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float kf_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float kf_sqrt(float x) { return __builtin_sqrtf(x); }

// 1/sqrt(x) in binary64 from the binary32 estimate and one Newton step
// (relative error ~3 e^2 with e ~ 1e-7 of the estimate).
__device__ __forceinline__ double kx_rsq(double x) {
  double y = (double)kf_rsq((float)x);
  return y * (1.5 - 0.5 * (x * (y * y)));
}

// 48-bit fixed-point offset with scale 2^(k+16): the high 32 bits as the
// returned word, the low 16 bits in lo16; binary32 fallback in the word (and
// lo16 = 0) when it does not fit.
__device__ __forceinline__ int kx_enc(double off, int k, bool& fallback, uint32_t& lo16) {
  // x = hi 2^16 + lo exactly (|x| < 2^47): hi = floor(x / 2^16), lo in [0, 2^16)
  const double x = rint(ldexp(off, k + 16));
  fallback = !(fabs(x) < 140737488000000.0);
  const double hi = floor(x * (1.0 / 65536.0));
  lo16 = fallback ? 0u : (uint32_t)fma(hi, -65536.0, x);
  return fallback ? __float_as_int((float)off) : (int)hi;
}
__device__ __forceinline__ double kx_dec(int w, uint32_t lo16, int k, bool fallback) {
  const double fx = ldexp(fma((double)w, 65536.0, (double)(lo16 & 0xFFFFu)), -(k + 16));   // exact
  const double ff = (double)__int_as_float(w);
  return fallback ? ff : fx;
}

// Element at byte offset `boff` of a per-tube base pointer: the base is uniform
// and the offset a 32-bit lane value, so loads and stores take the scalar-base
// + 32-bit offset addressing form (a tube segment holds < 2^28 slots).
template <typename T>
__device__ __forceinline__ T& kx_at(T* base, uint32_t boff) { return *(T*)((char*)base + boff); }

// Reference cell (1-based) of a cell-word config index qc: 0 -> 1, -1 -> nq.
KAOS_HD int kx_refq(int qc, int nq) { return (qc >= 1) ? qc : (qc == 0 ? 1 : nq); }

// Reference of the stored stage-4 q of an ion whose cell word names cell Q:
// the centre qGC of Q split into two binary32 parts, hi + lo (the push's LDS
// table holds the same two parts).  The state word holds the binary32 offset
// q - (hi + lo), so q is kept to ~1e-7 of its distance from the centre
// rather than of |q| (the field interpolation weight of a binary32 q would
// be off by ~1e-7 |q| / (cell spacing), ~1e-5 on a 256-cell line).
KAOS_HD double kx_qref(double qgc) {
  const float hi = (float)qgc;
  return (double)hi + (double)(float)(qgc - (double)hi);
}

// Parallel acceleration (signs as kaos_apar; qle0 = stage q <= 0).
__device__ __forceinline__ float kf_apar(bool qle0, float AG, float AEA, int grav, int eamb,
                                         int eambsign, bool with_grav) {
  KAOS_MIXED_CONTRACT
  float a = 0.0f;
  if (grav && with_grav) a = qle0 ? fabsf(AG) : -fabsf(AG);
  if (eamb) {
    if (eambsign) a = a + AEA;
    else a = a + (qle0 ? -fabsf(AEA) : fabsf(AEA));
  }
  return a;
}

// b-hat at (rho, z) off the line, as kaos_bhat_pt.
__device__ __forceinline__ void kf_bhat_pt(float rho, float z, float& bR, float& bZ) {
  KAOS_MIXED_CONTRACT
  float z2 = z * z;
  float u2 = rho * rho + z2;
  float d = kf_rsq(u2 * (u2 + 3.0f * z2));
  bR = (3.0f * (z * rho)) * d;
  bZ = (2.0f * z2 - rho * rho) * d;
}

// Cell of |r| on ascending binary64 radius edges (rGH[i] = |r| of qGH of cell
// i+1, rGL0 of qGL of cell 1), walked from the previous cell: 1..nq, 0 below
// the lower boundary (or NaN), -1 above the upper one.
__device__ __forceinline__ int kf_find_rcell_from(double r, const double* rGH, double rGL0,
                                                  double rGHlast, int nq, int hint) {
  if (!(r >= rGL0)) return 0;
  if (r > rGHlast) return -1;
  int i = hint - 1;
  if (i < 0) i = 0;
  if (i > nq - 1) i = nq - 1;
  while (i > 0 && r <= rGH[i - 1]) --i;
  while (!(r <= rGH[i])) ++i;
  return i + 1;
}

// q - qX for the centre row X = (hi, EA, EG, lo) of an ion stored as the
// offset `off` from row R: the hi parts are exact binary32 numbers, so each
// difference is rounded relative to itself, not to |q|.
__device__ __forceinline__ float kf_dq(const float4& R, float off, const float4& X) {
  KAOS_MIXED_CONTRACT
  return (R.x - X.x) + ((R.w - X.w) + off);
}

// Field interpolation at q = (R.hi + R.lo) + off, R = cE[hint - 1] (as
// kaos_interp2_from: the grid-edge test on q in binary64, the walk and the
// weight on the differences above).  Cell centres packed as (hi, EA, EG, lo).
__device__ __forceinline__ int kf_interp2_from(float off, const float4* cE, int nq, double qGL0, double qGHlast,
                                               int desc, int hint, float& out, float& out2) {
  KAOS_MIXED_CONTRACT
  const float4 R = cE[hint - 1];
  const double q = ((double)R.x + (double)R.w) + (double)off;
  int ub = nq - 1;
  int lo = hint - 1;
  if (lo < 1) lo = 1;
  if (lo > ub) lo = ub;
  if (desc) {
    if (!(q <= qGL0) || !(q >= qGHlast)) return 0;
    if (kf_dq(R, off, cE[0]) >= 0.0f) { out = cE[0].y; out2 = cE[0].z; return 1; }
    if (kf_dq(R, off, cE[ub]) <= 0.0f) { out = cE[ub].y; out2 = cE[ub].z; return 1; }
    while (lo > 1 && kf_dq(R, off, cE[lo - 1]) >= 0.0f) --lo;
    while (!(kf_dq(R, off, cE[lo]) >= 0.0f)) ++lo;
  } else {
    if (!(q >= qGL0) || !(q <= qGHlast)) return 0;
    if (kf_dq(R, off, cE[0]) <= 0.0f) { out = cE[0].y; out2 = cE[0].z; return 1; }
    if (kf_dq(R, off, cE[ub]) >= 0.0f) { out = cE[ub].y; out2 = cE[ub].z; return 1; }
    while (lo > 1 && kf_dq(R, off, cE[lo - 1]) <= 0.0f) --lo;
    while (!(kf_dq(R, off, cE[lo]) <= 0.0f)) ++lo;
  }
  const float4 A = cE[lo - 1], B = cE[lo];
  const float f = kf_dq(R, off, A) * kf_rcp((B.x - A.x) + (B.w - A.w));
  out = A.y + f * (B.y - A.y);
  out2 = A.z + f * (B.z - A.z);
  return 1;
}

// Linear pieces of the field along the tube for the common case: piece k
// (0..nq) runs from cell centre k-1 to centre k (0-based), piece 0 from the
// lower grid edge to centre 0 and piece nq from the last centre to the upper
// edge (constant there, as kaos_interp2_from).  In the direction of
// increasing cell index (sigma = -1 when q decreases with the index) an ion
// of cell h whose stored offset from centre h-1 is off lies in piece h when
// sigma off >= 0 (t = sigma off from centre h-1), else in piece h-1
// (t = sigma off + W from centre h-2); it is in the piece when 0 <= t <= W,
// W the piece's length, and the fields are (E, G) + t (dE, dG), already
// multiplied by qs/ms.  Any other ion takes kf_interp2_from.
__device__ __forceinline__ void kf_piece_set(float4* pe, float* pw, int k, int nq, const double* qGC, const double* EA,
                                             const double* EG, double qGL0, double qGHlast, int desc, double qsms) {
  const double sg = desc ? -1.0 : 1.0;
  const int b = (k > 0) ? k - 1 : 0;
  double w, dE = 0.0, dG = 0.0;
  if (k == 0) w = sg * (qGC[0] - qGL0);
  else if (k == nq) w = sg * (qGHlast - qGC[nq - 1]);
  else {
    w = sg * (qGC[k] - qGC[k - 1]);
    dE = qsms * (EA[k] - EA[k - 1]) / w;
    dG = qsms * (EG[k] - EG[k - 1]) / w;
  }
  pe[k] = make_float4((float)(qsms * EA[b]), (float)(qsms * EG[b]), (float)dE, (float)dG);
  pw[k] = (float)w;
}

// Encode the velocity state v (binary64) given the new Vparind ipn (0 = off grid).
__device__ __forceinline__ int kx_enc_v(double v, int ipn, const double* vGH, const int* vk, int nv,
                                        bool& vflag, uint32_t& lo16) {
  if (ipn >= 1) {
    double lo = (ipn == 1) ? vGH[nv] : vGH[ipn - 2];
    double ref = 0.5 * (lo + vGH[ipn - 1]);
    return kx_enc(v - ref, vk[ipn - 1], vflag, lo16);
  }
  vflag = true;
  lo16 = 0u;
  return __float_as_int((float)(v - ((v > vGH[nv - 1]) ? vGH[nv - 1] : vGH[nv])));
}
__device__ __forceinline__ double kx_dec_v(int w, uint32_t lo16, int ip, bool vflag, const double* vGH, const int* vk,
                                           int nv) {
  if (vflag) {
    double off = (double)__int_as_float(w);
    return ((off > 0.0) ? vGH[nv - 1] : vGH[nv]) + off;
  }
  double lo = (ip == 1) ? vGH[nv] : vGH[ip - 2];
  double ref = 0.5 * (lo + vGH[ip - 1]);
  return ref + kx_dec(w, lo16, vk[ip - 1], false);
}

// Binary32 constants of a tube's heating, set once per tube in LDS (the
// kick reads them instead of forming them from the binary64 parameters per
// ion): qs/ms, 1/OmegaG0, 0.5 (qs/ms)^2 EtaLH S0, XiPerp1/2, -ChiPerp1/2,
// lambdaPerp, the heated arc length and h.
struct kf_icr_c {
  float qm, iog0, c, xi1, xi2, nchi1, nchi2, lam, sumds, h;
};
__host__ __device__ inline kf_icr_c kf_icr_consts(const kaos_icr_t& P) {
  kf_icr_c k;
  const float qm = (float)(P.qs / P.ms);
  k.qm = qm;
  k.iog0 = (float)(1.0 / P.omega_g0);
  k.c = 0.5f * (qm * qm) * (float)(P.eta_lh * P.s0);
  k.xi1 = (float)P.xi1; k.xi2 = (float)P.xi2;
  k.nchi1 = -(float)P.chi1; k.nchi2 = -(float)P.chi2;
  k.lam = (float)P.lambda_perp;
  k.sumds = (float)P.sum_ds;
  k.h = (float)P.h;
  return k;
}

// Stochastic ICR kick of one ion in binary32 (the mixed form's heating; the
// arithmetic of kaos_icr_kick with hardware log2/exp2/sqrt/rcp and the
// sine/cosine of revolutions).  u0, u1 in (0, 1]: one Box-Muller pair.
// Products are ordered so that no intermediate leaves the binary32 range
// (qs/ms first: qs^2 and ms^2 alone underflow).
__device__ __forceinline__ void kf_icr_kick(const kf_icr_c& P, float r, float w, float vpar, float v1, float v2,
                                            float u0, float u1, float& v1N, float& v2N) {
  KAOS_MIXED_CONTRACT
  const float R = kf_sqrt(-2.0f * 0.69314718f * __builtin_amdgcn_logf(u0));
  const float g1 = R * __builtin_amdgcn_cosf(u1), g2 = R * __builtin_amdgcn_sinf(u1);
  const float qm = P.qm;
  const float B = (float)KAOS_MDIP * (kf_sqrt(4.0f - 3.0f * w) * kf_rcp((r * r) * r));
  const float og = qm * B;
  const float vperp = kf_sqrt(v1 * v1 + v2 * v2);
  float isig3 = 1.0f;
  if (P.lam != 0.0f && ((float)KAOS_TWOPI * vperp) * kf_rcp(og) >= P.lam) {
    const float sig = ((float)KAOS_TWOPI * vperp) * kf_rcp(P.lam * og);
    isig3 = kf_rcp((sig * sig) * sig);
  }
  const float lx = __builtin_amdgcn_logf(og * P.iog0);
  const float D1 = (P.c * P.xi1) * __builtin_amdgcn_exp2f(P.nchi1 * lx) * isig3;
  const float D2 = (P.c * P.xi2) * __builtin_amdgcn_exp2f(P.nchi2 * lx) * isig3;
  const float eps = fabsf(vpar) * kf_rcp(og * P.sumds);
  const float tau = (float)KAOS_TWOPI * kf_rcp(og * kf_sqrt(eps));
  const float h = P.h;
  float dv1 = 0.0f, dv2 = 0.0f;
  if (h <= tau) {
    const float f = h * kf_rcp(tau);
    dv1 = (f * kf_sqrt((2.0f * D1) * tau)) * g1;
    dv2 = (f * kf_sqrt((2.0f * D2) * tau)) * g2;
  }
  v1N = v1 + dv1;
  v2N = v2 + dv2;
}

// Push + bin, mixed form.  Same step semantics as k_push_fp64 (removal at a
// master step, pre-push binning, stale 2-D velocity index, stage-4 gravity
// gate, field interpolation at the previous stage-4 q on master steps).
// Static LDS (grids of up to MAXQ cells, so the arrays sit at fixed offsets):
// binary64 rGH, rref, vparGH + GL0; qGC as two binary32 parts (kx_qref),
// binary32 EA, EG; int rk, vk.
// Dynamic LDS: the histogram window (a.hwords words).  Uniform values are
// kept few (flag words, 32-bit offsets, h-multiples derived per ion).
// (launch bounds: minimum resident wavefronts per SIMD, which caps the VGPRs;
// 4 leaves the per-ion loop without VGPR spills)
#ifndef KAOS_MIXED_WAVES_EU
#define KAOS_MIXED_WAVES_EU 4
#endif
// ICR: the instantiation launched when a tube heats or has charge exchange
// (the plain one carries neither's registers).
// BLOCK: threads per block; each thread pushes KAOS_TILE / BLOCK ions of a
// tile (the autotuner times 256 and 512 for the plain instantiation).
template <int MAXQ, bool ICR, int BLOCK = KAOS_MIXED_BLOCK>
__global__ void __launch_bounds__(BLOCK, KAOS_MIXED_WAVES_EU) k_push_mixed(StepArgs a) {
  KAOS_MIXED_CONTRACT
  constexpr int BLK = BLOCK, PPT = KAOS_TILE / BLOCK;
  __shared__ double s_rGH[MAXQ], s_rref[MAXQ], s_vGH[KAOS_MAXV + 1];
  __shared__ float4 s_cE[MAXQ];
  __shared__ float4 s_pe[MAXQ + 1];
  __shared__ float s_pw[MAXQ + 1];
  __shared__ int s_rk[MAXQ], s_vk[KAOS_MAXV + 1];
  extern __shared__ unsigned int s_hist[];
  __shared__ unsigned int s_pushed;
  __shared__ int s_wq0, s_wnq, s_wtube, s_group;
  __shared__ kf_icr_c s_icc;                     // heating constants of the current tube (ICR instantiation)
  __shared__ int4 s_wbox;
  const int tid = threadIdx.x;
  int cur = -1;
  const int sflags = (a.do_interp ? 1 : 0) | (a.do_count ? 2 : 0) | (a.do_remove ? 4 : 0) | (a.do_qk ? 8 : 0);
  // per-tube scalars and base pointers
  int hoff = 0, goff = 0;
  long long toff = 0;
  uint32_t tcount = 0;
  uint32_t* cellT = nullptr;
  int *rdiT = nullptr, *vpiT = nullptr;
  uint32_t* loT = nullptr;
  float *qkfT = nullptr, *aefT = nullptr, *agfT = nullptr;
  int nq = 0, nvpar = 0, nvperp = 0;
  int tflags = 0;                      // desc | grav << 1 | eamb << 2 | eambsign << 3 | (rs < 0) << 4
  double rGL0 = 0.0, rGHlast = 0.0;
  double qGL0 = 0.0, qGHlast = 0.0;
  float qsms = 0.0f, vscale = 0.0f;
  double ip = 0.0, ip2 = 0.0, hd = 0.0;
  uint32_t tbox = 0u;
  int icr = 0;                         // ICR heating in this tube (ICR instantiation only)
  int tcx = 0;                         // charge exchange in this tube (cos(thetak4) kept)
  int woff = 0;                        // the tube's Vperp edges in a.vperpGH
  for (int pass = 0;; ++pass) {
  const int g = kaos_next_group(a, pass, &s_group);
  if (g < 0) break;
  int t0 = g * a.tiles_per_block;
  int t1 = min(t0 + a.tiles_per_block, a.ntiles);
  unsigned int npushed = 0;
  if (tid == 0) {
    s_pushed = 0u;
    int tb0 = (t0 < t1) ? a.tile_tube[t0] : 0;
    const DevTube& T0 = a.tubes[tb0];
    const int4 bx = a.vbox[tb0];
    int wnq = (sflags & 2) ? (2 * a.hwords) / (bx.y * bx.w) : 0;
    if (wnq > T0.nq) wnq = T0.nq;
    int q0 = 1;
    if (t0 < t1 && wnq > 0) q0 = kaos_window_q0(a, t0, t1, tb0, T0.nq, T0.off + a.count[tb0], wnq);
    s_wq0 = q0; s_wnq = wnq; s_wtube = tb0; s_wbox = bx;
  }
  for (int i = tid; i < a.hwords; i += BLK) s_hist[i] = 0u;
  __syncthreads();
  const int wq0 = s_wq0, wnq = s_wnq, wtube = s_wtube;
  const int4 wb = s_wbox;
  for (int tile = t0; tile < t1; ++tile) {
    int tb = a.tile_tube[tile];
    if (tb != cur) {
      __syncthreads();
      const DevTube& T = a.tubes[tb];
      nq = T.nq; nvpar = T.nvpar; nvperp = T.nvperp;
      tflags = (T.desc ? 1 : 0) | (T.grav ? 2 : 0) | (T.eamb ? 4 : 0) | (T.eambsign ? 8 : 0) |
               (T.rsgn < 0.0 ? 16 : 0);
      goff = (int)T.goff; hoff = (int)T.hoff;
      qGL0 = T.qGL0; qGHlast = T.qGHlast; rGL0 = T.rGL0; rGHlast = T.rGHlast;
      qsms = (float)T.qsms;
      hd = T.h;
      ip = 1.0 / T.p;
      ip2 = ip * ip;
      for (int i = tid; i < nq; i += BLK) {
        s_rGH[i] = a.rGH[goff + i]; s_rref[i] = a.rref[goff + i];
      {
        const double qg = a.qGC[goff + i];
        const float qh = (float)qg;
        s_cE[i] = make_float4(qh, (float)a.EA[goff + i], (float)a.EG[goff + i], (float)(qg - (double)qh));
      }
        s_rk[i] = a.rk[goff + i];
      }
      for (int k = tid; k <= nq; k += BLK)
        kf_piece_set(s_pe, s_pw, k, nq, a.qGC + goff, a.EA + goff, a.EG + goff, T.qGL0, T.qGHlast, T.desc, T.qsms);
      for (int i = tid; i <= nvpar; i += BLK) s_vGH[i] = a.vparGH[T.voff + i];
      for (int i = tid; i < nvpar; i += BLK) s_vk[i] = a.vk[T.voff + i];
      toff = T.off;
      tcount = (uint32_t)a.count[tb];
      cellT = a.cell + toff; rdiT = a.rdi + toff; vpiT = a.vpi + toff; loT = a.lo + toff;
      qkfT = a.qkf + toff; aefT = a.aef + toff; agfT = a.agf + toff;
      vscale = (float)((double)nvpar / (a.vparGH[T.voff + nvpar - 1] - a.vparGH[T.voff + nvpar]));
      tbox = kaos_pack_box(a.vbox[tb]);
      if (ICR) {
        icr = T.icr;
        tcx = T.cx;
        woff = (int)T.woff;
        if (tid == 0 && icr) s_icc = kf_icr_consts(T.icrp);
      }
      cur = tb;
      __syncthreads();
    }
    const uint32_t start = (uint32_t)(a.tile_start[tile] - toff);
#pragma unroll 1
    for (int it = 0; it < PPT; ++it) {
      const uint32_t jl = start + (uint32_t)(it * BLK + tid);
      if (jl >= tcount) break;
      const uint32_t b4 = jl * 4u;
      const uint32_t cold = kx_at(cellT, b4), wl = kx_at(loT, b4);
      const int wr = kx_at(rdiT, b4), wv = kx_at(vpiT, b4);
      const float wq = kx_at(qkfT, b4);
      if (cold == KAOS_DEAD) continue;
      int qprev = (int)((cold >> 16) & KAOS_QMASK) - 1;
      if ((sflags & 4) && (qprev == 0 || qprev == -1)) {
        kx_at(cellT, b4) = KAOS_DEAD;
        atomicAdd(&a.esc_acc[2 * cur + (qprev == 0 ? 0 : 1)], 1ull);
        continue;
      }
      ++npushed;
      const int hint = kx_refq(qprev, nq);
      const int ipc = (int)(cold & 0xFFu), iq = (int)((cold >> 8) & 0xFFu);
      // decode, with the hint cell's |r| edges read alongside (no loop: the
      // search below walks only when the ion has left the hint cell)
      const double rlo_h = (hint > 1) ? s_rGH[hint - 2] : rGL0, rhi_h = s_rGH[hint - 1];
      const int ipv = max(ipc, 1);
      const bool rfl = (cold & KAOS_RFLAG) != 0, vfl = (cold & KAOS_VFLAG) != 0;
      const double rs = s_rref[hint - 1] + kx_dec(wr, wl, s_rk[hint - 1], rfl);
      const double voff = (double)__int_as_float(wv);
      const double vs64 =
          vfl ? (((voff > 0.0) ? s_vGH[nvpar - 1] : s_vGH[nvpar]) + voff)
              : 0.5 * (((ipv == 1) ? s_vGH[nvpar] : s_vGH[ipv - 2]) + s_vGH[ipv - 1]) +
                    kx_dec(wv, wl >> 16, s_vk[ipv - 1], false);
      float aea, ag;
      const int eambsign = (tflags >> 3) & 1, grav = (tflags >> 1) & 1, eamb = (tflags >> 2) & 1;
      if (sflags & 1) {
        // fields at the stored stage-4 q: the hint cell's piece, else the walk
        const float so = (tflags & 1) ? -wq : wq;
        const bool up = so >= 0.0f;
        const int k = up ? hint : hint - 1;
        const float4 pe = s_pe[k];
        const float pw = s_pw[k];
        const float t = up ? so : so + pw;
        float ea = 0.0f, eg = 0.0f;
        bool ok = (t >= 0.0f) && (t <= pw);
        if (ok) {
          ea = fmaf(t, pe.z, pe.x);
          eg = fmaf(t, pe.w, pe.y);
        } else {
          ok = kf_interp2_from(wq, s_cE, nq, qGL0, qGHlast, tflags & 1, hint, ea, eg);
          ea *= qsms; eg *= qsms;
        }
        if (ok) {
          // AEAmag = qs/ms E (|.| unless EAMBSIGN), AGmag = |qs/ms G|
          aea = eambsign ? ea : fabsf(ea);
          ag = fabsf(eg);
          kx_at(aefT, b4) = aea; kx_at(agfT, b4) = ag;
        } else {
          aea = kx_at(aefT, b4); ag = kx_at(agfT, b4);
        }
      } else {
        aea = kx_at(aefT, b4); ag = kx_at(agfT, b4);
      }
      // stage A on the line: cell from |r| (binary64); sin, cos, b-hat to ~1e-14
      const float h = (float)hd, h2 = 0.5f * h, hsq2 = 0.5f * (h * h), h6 = h * (1.0f / 6.0f),
                  hh6 = (h * h) * (1.0f / 6.0f);
      const double r = fabs(rs);
      const double irep = ip * (1.0 / KAOS_RE);
      const double w = r * irep;
      const double omw = 1.0 - w;
      int qc = hint;
      {
        const bool other_hemi = (rs < 0.0) != ((tflags & 16) != 0);
        const bool in_hint = ((hint > 1) ? (r > rlo_h) : (r >= rlo_h)) && r <= rhi_h;
        if (other_hemi || !in_hint) qc = other_hemi ? -1 : kf_find_rcell_from(r, s_rGH, rGL0, rGHlast, nq, hint);
      }
      const double rw = kx_rsq(w);
      const double s64 = w * rw;
      double c64 = omw * kx_rsq(omw);
      if (rs < 0.0) c64 = -c64;
      const double isl64 = kx_rsq(4.0 - 3.0 * w);
      const double bR64 = (3.0 * (c64 * s64)) * isl64, bZ64 = (2.0 - 3.0 * w) * isl64;
      const double rho64 = r * s64, z64 = r * c64;
      const float vs = (float)vs64;
      const float cf = (float)c64;
      const float rho = (float)rho64, z = (float)z64;
      const float bR = (float)bR64, bZ = (float)bZ64;
      const float ap1 = kf_apar(cf <= 0.0f, ag, aea, grav, eamb, eambsign, true);
      const float a1R = ap1 * bR, a1Z = ap1 * bZ;
      const float VRR = vs * bR, VRZ = vs * bZ;
      // stage B at the half step
      float b2R, b2Z;
      const float zB = z + h2 * VRZ;
      kf_bhat_pt(rho + h2 * VRR, zB, b2R, b2Z);
      const float ap2 = kf_apar(zB <= 0.0f, ag, aea, grav, eamb, eambsign, true);
      const float a2R = ap2 * b2R, a2Z = ap2 * b2Z;
      // stage C at the full-step predictor
      float b4R, b4Z;
      const float dRC = h * VRR + hsq2 * a1R, dZC = h * VRZ + hsq2 * a1Z;
      const float zC = z + dZC;
      kf_bhat_pt(rho + dRC, zC, b4R, b4Z);
      // charge exchange (heating/charge-exchange instantiation): cos(thetak4),
      // the frame of a conversion at the next step
      if (ICR && tcx) {
        const float rC = rho + dRC;
        kx_at(a.ck4f + toff, b4) = zC * kf_rsq(rC * rC + zC * zC);
      }
      const float ap4 = kf_apar(zC <= 0.0f, ag, aea, grav, eamb, eambsign, qc != 0 && qc != -1);
      const float a4R = ap4 * b4R, a4Z = ap4 * b4Z;
      // RK4 (k2 = k3): dV = h (a1 + 4 a2 + a4)/6, dx = h V + h^2 (2 a1 + a2)/6
      const float dVR = h6 * (((a1R + 2.0f * a2R) + 2.0f * a2R) + a4R);
      const float dVZ = h6 * (((a1Z + 2.0f * a2Z) + 2.0f * a2Z) + a4Z);
      const double hv = hd * vs64;
      const double dR = hv * bR64 + (double)(hh6 * (2.0f * a1R + a2R));
      const double dZ = hv * bZ64 + (double)(hh6 * (2.0f * a1Z + a2Z));
      // new point and root solve: u2 = r^2 + Tu, Z^2 = r^2 (1 - w) + Tz
      const double Tu = 2.0 * (rho64 * dR + z64 * dZ) + (dR * dR + dZ * dZ);
      const double Tz = 2.0 * (z64 * dZ) + dZ * dZ;
      const double r2 = r * r;
      const double u2N = r2 + Tu;
      const double Z2 = r2 * omw + Tz;
      const double g0 = (r + (s64 * dR + c64 * dZ)) * (1.0 / KAOS_RE);
      const double rN = kaos_quartic_newton((KAOS_RE2 * KAOS_RE2) * Z2, (u2N * u2N) * u2N, ip, g0) * KAOS_RE;
      const double rsN = (z64 + dZ <= 0.0) ? -rN : rN;
      // V|| at the new point
      const float wN = (float)(rN * irep);
      float csN = kf_sqrt(wN * (1.0f - wN));
      if (rsN < 0.0) csN = -csN;
      const float islN = kf_rsq(4.0f - 3.0f * wN);
      const float bNR = (3.0f * csN) * islN, bNZ = (2.0f - 3.0f * wN) * islN;
      const float eR = bNR - bR, eZ = bNZ - bZ;
      const double vN = (vs64 - vs64 * (double)(0.5f * (eR * eR + eZ * eZ))) + (double)(bNR * dVR + bNZ * dVZ);
      // binning of the pre-push state: Vparind from the cell word
      if ((sflags & 2) && qc >= 1) {
        if (ipc >= 1 && iq >= 1) {
          const int wql = qc - wq0, jr = ipc - wb.x, ir = iq - wb.z;
          if (cur == wtube && (unsigned)wql < (unsigned)wnq && (unsigned)jr < (unsigned)wb.y &&
              (unsigned)ir < (unsigned)wb.w) {
            const int li = (wql * wb.w + ir) * wb.y + jr;
            atomicAdd(&s_hist[li >> 1], 1u << (16 * (li & 1)));
          } else {
            atomicAdd(&a.hist[hoff + ((long long)(qc - 1) * nvperp + (iq - 1)) * nvpar + (ipc - 1)], 1u);
            kaos_hbox_note(&a.hbox[goff + qc - 1], tbox, ipc, iq);
          }
        } else {
          atomicAdd(&a.nqoff[goff + qc - 1], 1u);
        }
      }
      if (rsN != rsN || vN != vN) set_status(a.status, 1u << KAOS_E_NAN);
      // stage-4 q = RE^2 z / r^3 at the full-step predictor
      // as the offset from the reference of the new cell word's cell:
      // q_C - q = (RE/r)^2 [b S + c (S - 1)], S = (1 + x)^(-3/2),
      // x = 2 (s a + c b) + a^2 + b^2, (a, b) = (dR, dZ)_C / r, with
      // S - 1 = -3/2 x + 15/8 x^2 - 35/16 x^3 (the next term, 2.46 x^4, is
      // below 1e-8 of S - 1 for |x| < 2e-3, a step of v dt < 6 km): binary32
      // terms that are small against q, each to ~1e-7 of itself
      if (sflags & 8) {
        const float ir = kf_rcp((float)r);
        const float ac = dRC * ir, bc = dZC * ir;
        const float sf = (float)s64;
        const float x = 2.0f * (sf * ac + cf * bc) + (ac * ac + bc * bc);
        const float Sm1 = x * fmaf(x, fmaf(x, -2.1875f, 1.875f), -1.5f);
        const float dq = bc + (bc + cf) * Sm1;
        const double rw2 = rw * rw;
        const double K = ip2 * (rw2 * rw2);                      // (RE / r)^2
        const float4 R = s_cE[kx_refq(qc, nq) - 1];
        kx_at(qkfT, b4) = (float)(fma(K, (double)dq, c64 * K - ((double)R.x + (double)R.w)));
      }
      // next push's Vparind (stale when off the grid) and the new state words
      // (the cell of the stored V|| first, then the general search)
      int ipn = ipc;
      {
        const double vlo = (ipv == 1) ? s_vGH[nvpar] : s_vGH[ipv - 2], vhi = s_vGH[ipv - 1];
        if (!(ipc >= 1 && ((ipc > 1) ? (vN > vlo) : (vN >= vlo)) && vN <= vhi))
          ipn = kaos_find_vcell_guess(vN, s_vGH, s_vGH[nvpar], nvpar, (double)vscale);
      }
      bool rflag, vflag;
      uint32_t rlo, vlo = 0u;
      const int Qn = kx_refq(qc, nq);
      kx_at(rdiT, b4) = kx_enc(rsN - s_rref[Qn - 1], s_rk[Qn - 1], rflag, rlo);
      if (ipn >= 1) {
        const double vref = 0.5 * (((ipn == 1) ? s_vGH[nvpar] : s_vGH[ipn - 2]) + s_vGH[ipn - 1]);
        kx_at(vpiT, b4) = kx_enc(vN - vref, s_vk[ipn - 1], vflag, vlo);
      } else {
        vflag = true;
        kx_at(vpiT, b4) = __float_as_int((float)(vN - ((vN > s_vGH[nvpar - 1]) ? s_vGH[nvpar - 1] : s_vGH[nvpar])));
      }
      kx_at(loT, b4) = rlo | (vlo << 16);
      // ICR heating (WaveHeatingSub at stage A, _8_1_1_KineticUpdateA.f90:1392) of
      // an ion inside the grid: binary32 kick (kf_icr_kick) from the Philox4x32
      // stream; the cell word then carries the Vperp cell of the kicked speed,
      // the next push's binning cell (off the grid the previous one stays)
      int iqw = iq;
      if (ICR && icr && qc != 0 && qc != -1) {
        const float v1 = kx_at(a.vrf + toff, b4), v2 = kx_at(a.vrf2 + toff, b4);
        uint32_t key[2] = { (uint32_t)a.seed, (uint32_t)(a.seed >> 32) ^ 0x49435248u };
        const long long idj = kx_at(a.id + toff, 2u * b4);
        uint32_t ctr[4] = { (uint32_t)idj, (uint32_t)(*a.stepno), (uint32_t)cur,
                            0x49435230u ^ (uint32_t)((unsigned long long)idj >> 32) };
        uint32_t o4[4];
        kaos_philox4x32(ctr, key, o4);
        const float u0 = (float)(((double)o4[0] + 1.0) * (1.0 / 4294967296.0));
        const float u1 = (float)(((double)o4[1] + 1.0) * (1.0 / 4294967296.0));
        float v1N, v2N;
        kf_icr_kick(s_icc, (float)r, (float)w, vs, v1, v2, u0, u1, v1N, v2N);
        kx_at(a.vrf + toff, b4) = v1N;
        kx_at(a.vrf2 + toff, b4) = v2N;
        const double* GH = a.vperpGH + woff;
        const int iqn = kaos_find_vcell_from((double)kf_sqrt(v1N * v1N + v2N * v2N), GH, GH[nvperp], nvperp, iq);
        if (iqn >= 1) iqw = iqn;
      }
      kx_at(cellT, b4) = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iqw << 8) | (uint32_t)(ipn >= 1 ? ipn : ipc) |
                  (rflag ? KAOS_RFLAG : 0u) | (vflag ? KAOS_VFLAG : 0u);
    }
  }
  __syncthreads();
  atomicAdd(&s_pushed, npushed);
  if (wnq > 0) {
    const DevTube& TW = a.tubes[wtube];
    kaos_window_flush(s_hist, a.hist, TW.hoff, TW.nvperp, TW.nvpar, wq0, wnq, wb, tid >> 6, BLK >> 6, tid & 63);
  }
  __syncthreads();
  if (tid == 0 && s_pushed) atomicAdd(a.pushed, (unsigned long long)s_pushed);
  }
}
