// kaos_hip.hip -- gfx950 kernels and C-ABI of libkaos_hip.so: the KAOS
// per-timestep kinetic hot path (push -> bin -> moments -> field -> boundary)
// on one MI355X.
//
// Particle layout (structure of arrays in HBM, one segment per flux tube):
//   rs    double  signed radius of the ion's field-line point (0 = dead slot)
//   vp    double  parallel velocity V||; the Cartesian velocity is V|| b-hat
//   vr    double  perpendicular speed Vperp (constant without wave heating)
//   qk    double  stage-4 dipole q of the last push before a master step
//   ae    double  AEAmag, ag double AGmag (per-ion field accelerations)
//   cell  uint32  (Qindk1+1) << 16 | Vperpind << 8 | Vparind of the last push
// (rs, vp) determine the reference's Cartesian state (x,y,z,Vx,Vy,Vz) exactly:
// the push ends every step with xN = pos(rs), VN = VparNp b-hat(xN), so the
// six Cartesian doubles are recomputed bit for bit from the two stored ones.
//
// Ions whose last push found them outside the domain are marked dead (rs = 0)
// at the next master step, which is when BoundaryConditionsSub removes them;
// dead slots are skipped, live ions keep their order, injected ions are
// appended, so the live ions read in slot order are the reference's packed
// arrays (BCReIndexSub, _8_0_7_1_BCReIndex.f90:33-133).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "kaos_dipole.h"
#include "../../include/kaos_hip.h"

#define KAOS_BLOCK 256                       // threads of the generic per-ion kernels
#define KAOS_TILE 1024                       // ions per tile
#define KAOS_TILE_PPT (KAOS_TILE / KAOS_BLOCK)
// Threads of the push blocks (256-thread blocks measured faster than 320 at
// the same wavefronts per SIMD)
#define KAOS_MIXED_BLOCK 256
#define KAOS_FP64_BLOCK 256
#define KAOS_PUSH_BLOCK_DEFAULT 256          // push threads per block: 256 (4 ions each) or 512 (2 each)
#define KAOS_MAXQ 512
#define KAOS_MAXV 255
#ifndef KAOS_HWORDS
#define KAOS_HWORDS 6144                     // LDS histogram window: 12288 16-bit counters
#endif
#define KAOS_TPB 16                          // tiles per push group (<= 63: 16-bit counters)
// Master steps between reorders by default (200-step averages on the 1e8-ion
// tube: the push keeps its speed for ~100 steps after a reorder once its LDS
// window spans the tube's velocity box, so reorders are rare).
#define KAOS_REORDER_FP64 128
#define KAOS_REORDER_MIXED 128
// Timed launches per autotuner candidate (its time is the fastest of them: a
// single sample lets the step-to-step spread, a few percent, pick the shape)
#define KAOS_TUNE_SAMPLES 3

struct DevTube {
  int nq, nvpar, nvperp, desc;
  int grav, eamb, eambsign, filter_pt;
  int lbcond, ubcond, south, filter_pt_m1;
  long long off, cap, n_inj_lb, n_inj_ub;
  long long goff, voff, woff, hoff, doff;    // grid / vpar / vperp / histogram / d3v offsets
  double p, cphi, sphi, h, ms, qs, qsms, nsnormfac, sig_lb, sig_ub;
  double qGL0, qGHlast, qGL_last0;
  double rGL0, rGHlast, rsgn;               // |r| of the outer q edges; sign of rs in the tube
  long long koff;                           // offset of the tube's nq + 2 reorder keys
  int icr;                                  // ICR wave heating on; its parameters (icrp)
  int ea_in;                                // KAOS_FIELD_*: source of EAmagRT at master steps
  kaos_icr_t icrp;
  int nv1, nv2;                             // 3-D velocity grid: Vperp1 / Vperp2 cells (0: 2-D grid)
  long long v1off, v2off, h3off, d3off;     // Vperp1 / Vperp2 edges, 3-D counts, d3v3 offsets
  int cx, cx_filter, cx_filter_pt, cx_pad;  // charge exchange: on, MOMENTFILTERflagT, M1Perp/M2Par window
  double sig_n, sig_scale;                  // neutral thermal speed sqrt(kB TNeut / mNeut); cross-section factor
  long long eoff, ecap;                     // the tube's ENA slots [eoff, eoff + ecap)
  int nvp, nvq, nvphi, ena_pad;             // ENA velocity grid cells
  long long epoff, eqoff, efoff;            // offsets of its Vp / Vq / Vphi cell arrays
  long long eboff, ehoff;                   // metric arrays [Q][bin] (hh, d33v); counts [Q][bin]
};

struct StepArgs {
  const DevTube* tubes;
  const int* tile_tube;
  const long long* tile_start;
  int ntiles, tiles_per_block;
  // per-cell concatenated arrays
  const double *qGH, *qGC, *EA, *EG;
  const double *vparGH, *vperpGH;
  // particles
  double *rs, *vp, *vr, *qk, *ae, *ag;       // binary64 form (vr: Vperp1)
  double *vr2;                               // binary64 Vperp2
  int *rdi, *vpi;                            // mixed form: fixed-point radius / V|| offsets (bits 16..47)
  uint32_t* lo;                              // mixed form: bits 0..15 of both offsets (radius | V|| << 16)
  float *vrf, *qkf, *aef, *agf;              // mixed form: Vperp1, stage-4 q offset (from kx_qref), fields
  float *vrf2;                               // mixed form: Vperp2
  const double *rGH, *rref;                  // |r| of the qGH edges; signed r of the cell centres
  const int *rk, *vk;                        // fixed-point exponents per config / Vpar cell
  int lds_nq, lds_nv;                        // largest nq / nvpar over the tubes (LDS sizing)
  uint32_t* cell;
  uint16_t* c12;                             // mixed form, 3-D grid: Vperp1 | Vperp2 << 8 cells (binary64)
  long long* id;                             // order key: upload index, then injection order (64-bit: never wraps)
  const long long* count;
  unsigned int *hist, *nqoff;               // phase-space counts; config counts off the v-grid
  unsigned long long* esc_acc;
  unsigned int* status;
  unsigned long long* pushed;                // live ions pushed (particle-steps)
  int do_interp, do_count, do_remove, do_qk;
  const int4* vbox;                          // per tube: Vpar cells [x, x + y), Vperp cells [z, z + w)
  int4* hbox;                                // per (tube, cell): Vpar cells [x, y], Vperp cells [z, w] holding counts
  const long long* stepno;                   // device step counter (keys the heating stream)
  unsigned long long seed;
  // launch shape: tiles per group, LDS histogram window (32-bit words), persistent
  // blocks taking groups from the counter `work` (else one group per block)
  int hwords, persistent, ngroups;
  unsigned int* work;
  // 3-D velocity grid (null in the 2-D form): Vperp1 / Vperp2 high edges (GL0
  // appended), counts [Q][V1][V2][Vpar] per tube
  const double *v1GH, *v2GH;
  unsigned int* hist3;
  const int2* grp3;                          // per binning block: its tiles [x, y), all of one tube
  // charge exchange: cos(thetak4) of each ion's last push (null without it)
  double* ck4;                               // binary64 form
  float* ck4f;                               // mixed form
};

__device__ __forceinline__ void set_status(unsigned int* st, unsigned int bit) { atomicOr(st, bit); }

// Work distribution of the push kernels: group g is tiles [g tpb, g tpb + tpb).
// Grid launches give block b group b; persistent launches (a few blocks per
// CU) take groups from an atomic counter until none is left, so every block
// exits once the groups are drained.
__device__ __forceinline__ int kaos_next_group(const StepArgs& a, int pass, int* s_group) {
  if (!a.persistent) return (pass == 0 && (int)blockIdx.x < a.ngroups) ? (int)blockIdx.x : -1;
  __syncthreads();
  if (threadIdx.x == 0) *s_group = (int)atomicAdd(a.work, 1u);
  __syncthreads();
  int g = *s_group;
  return (g < a.ngroups) ? g : -1;
}

// A tube's velocity box packed in 32 bits: first Vpar cell | count << 8 |
// first Vperp cell << 16 | count << 24 (cell indices and counts <= 255).
__device__ __forceinline__ uint32_t kaos_pack_box(int4 b) {
  return (uint32_t)b.x | ((uint32_t)b.y << 8) | ((uint32_t)b.z << 16) | ((uint32_t)b.w << 24);
}

// Bounding box of a (tube, config cell)'s counts this master step (int4 as
// x0, x1, z0, z1, inclusive), set to the tube's velocity box before the push;
// a count added outside that box (packed in pb) widens it.  The histogram
// clear and the moment sums then touch only the box.
__device__ __forceinline__ void kaos_hbox_note(int4* hb, uint32_t pb, int ip, int iq) {
  if ((unsigned)(ip - (int)(pb & 0xFFu)) >= ((pb >> 8) & 0xFFu) ||
      (unsigned)(iq - (int)((pb >> 16) & 0xFFu)) >= (pb >> 24)) {
    int* h = (int*)hb;
    atomicMin(&h[0], ip); atomicMax(&h[1], ip);
    atomicMin(&h[2], iq); atomicMax(&h[3], iq);
  }
}

// First config cell of a push group's histogram window of wnq cells: centred
// on the cells (cell words of the previous push, 10-bit field qc + 1) of the
// group's first and last ions, which bracket the group's cells when the ions
// were last sorted.
__device__ __forceinline__ int kaos_window_q0(const StepArgs& a, int t0, int t1, int tb0, int nq, long long jend,
                                              int wnq) {
  const long long j0 = a.tile_start[t0];
  long long j1 = (a.tile_tube[t1 - 1] == tb0) ? a.tile_start[t1 - 1] + KAOS_TILE - 1 : jend - 1;
  if (j1 >= jend) j1 = jend - 1;
  int qa = 0, qb = 0;
  if (j0 < jend) qa = (int)((a.cell[j0] >> 16) & 0x3FFu) - 1;
  if (j1 >= j0 && j1 < jend) qb = (int)((a.cell[j1] >> 16) & 0x3FFu) - 1;
  if (qa < 1 || qa > nq) qa = 0;
  if (qb < 1 || qb > nq) qb = 0;
  const int qm = (qa && qb) ? (qa + qb) / 2 : (qa ? qa : (qb ? qb : 1));
  int q0 = qm - (wnq - 1) / 2;
  if (q0 + wnq - 1 > nq) q0 = nq - wnq + 1;
  if (q0 < 1) q0 = 1;
  return q0;
}

// Flush of a push group's LDS window (16-bit counters, two per word) to the
// global histogram.  The window holds config cells [wq0, wq0 + wnq) x the
// tube's velocity box b (Vpar cells [b.x, b.x + b.y), Vperp cells [b.z, b.z +
// b.w)), Vpar fastest; wavefront w of nw takes rows (cell, Vperp cell) w, w +
// nw, ..., its lanes the row's Vpar cells.
__device__ __forceinline__ void kaos_window_flush(const unsigned int* s_hist, unsigned int* hist, long long hoff,
                                                  int nvperp, int nvpar, int wq0, int wnq, int4 b, int w, int nw,
                                                  int lane) {
  const int nrows = wnq * b.w;
  for (int rr = w; rr < nrows; rr += nw) {
    const int wql = rr / b.w, ir = rr - wql * b.w;
    const long long g0 = hoff + ((long long)(wq0 - 1 + wql) * nvperp + (b.z - 1 + ir)) * nvpar + (b.x - 1);
    for (int jr = lane; jr < b.y; jr += 64) {
      const int li = rr * b.y + jr;
      const unsigned int v = (s_hist[li >> 1] >> (16 * (li & 1))) & 0xFFFFu;
      if (v) atomicAdd(&hist[g0 + jr], v);
    }
  }
}

#include "kaos_mixed.h"

// ---------------------------------------------------------------------------
// Push + bin (FP64 form).  One RK4 step per live ion (KineticRK4UpdateSub,
// _8_1_KineticRK4Update.f90:29-314) with, on master steps, the field
// interpolation at the previous stage-4 q (_8_7_AmbipolarEfield.f90:1453-2720,
// _8_7_2_Gravfield.f90:132-...) and the phase-space histogram of the
// pre-push cells (_8_3_1_IonParticleCounts.f90:612-634); at a master step,
// ions whose previous-push cell is 0/-1 are first marked dead
// (_8_0_7_BoundaryConditions.f90:158-256).

// Cell centres packed for the field interpolation: q, EA, EG.
struct __align__(16) KaosCE64 { double q, ea, eg, pad; };

// kaos_interp2_from on packed centres: the same walk, interval and formula.
__device__ __forceinline__ int kd_interp2_from(double q, const KaosCE64* cE, int nq, double qGL0, double qGHlast,
                                               int desc, int hint, double* out, double* out2) {
  int ub = nq - 1;
  int lo = hint - 1;
  if (lo < 1) lo = 1;
  if (lo > ub) lo = ub;
  if (desc) {
    if (!(q <= qGL0) || !(q >= qGHlast)) return 0;
    if (q >= cE[0].q) { *out = cE[0].ea; *out2 = cE[0].eg; return 1; }
    if (q <= cE[ub].q) { *out = cE[ub].ea; *out2 = cE[ub].eg; return 1; }
    while (lo > 1 && q >= cE[lo - 1].q) --lo;
    while (!(q >= cE[lo].q)) ++lo;
  } else {
    if (!(q >= qGL0) || !(q <= qGHlast)) return 0;
    if (q <= cE[0].q) { *out = cE[0].ea; *out2 = cE[0].eg; return 1; }
    if (q >= cE[ub].q) { *out = cE[ub].ea; *out2 = cE[ub].eg; return 1; }
    while (lo > 1 && q <= cE[lo - 1].q) --lo;
    while (!(q <= cE[lo].q)) ++lo;
  }
  const KaosCE64 A = cE[lo - 1], B = cE[lo];
  double f = (q - A.q) / (B.q - A.q);
  *out = A.ea + (f * (B.ea - A.ea));
  *out2 = A.eg + (f * (B.eg - A.eg));
  return 1;
}

// Common case first: q between the walk start's centre and one of its
// neighbours (three centres read together, no loop); else the walk.  Both
// give the same interval, so the same value.  Needs nq >= 2.
__device__ __forceinline__ int kd_interp2_fast(double q, const KaosCE64* cE, int nq, double qGL0, double qGHlast,
                                               int desc, int hint, double* out, double* out2) {
  const int ub = nq - 1;
  const int lo = max(1, min(hint - 1, ub));
  const KaosCE64 L = cE[lo - 1], M = cE[lo], H = cE[min(lo + 1, ub)];
  const bool up = desc ? !(q >= M.q) : !(q <= M.q);
  const bool ok = up ? ((desc ? (q >= H.q) : (q <= H.q)) && lo < ub) : (desc ? (q < L.q) : (q > L.q));
  if (ok) {
    const double x1 = up ? M.q : L.q, x2 = up ? H.q : M.q;
    const double e1 = up ? M.ea : L.ea, e2 = up ? H.ea : M.ea;
    const double g1 = up ? M.eg : L.eg, g2 = up ? H.eg : M.eg;
    double f = (q - x1) / (x2 - x1);
    *out = e1 + (f * (e2 - e1));
    *out2 = g1 + (f * (g2 - g1));
    return 1;
  }
  return kd_interp2_from(q, cE, nq, qGL0, qGHlast, desc, hint, out, out2);
}

// Config cell of q: the hint cell tested first (both of its edges read
// together), else kaos_find_qcell_from (same result).
__device__ __forceinline__ int kd_find_qcell_fast(double q, const double* qGH, double qGL0, double qGHlast, int nq,
                                                  int desc, int hint) {
  const int i = max(0, min(hint - 1, nq - 1));
  const double hi = qGH[i], lo = (i > 0) ? qGH[i - 1] : qGL0;
  const bool ok = desc ? (((i > 0) ? (q < lo) : (q <= lo)) && q >= hi)
                       : (((i > 0) ? (q > lo) : (q >= lo)) && q <= hi);
  return ok ? i + 1 : kaos_find_qcell_from(q, qGH, qGL0, qGHlast, nq, desc, hint);
}

// Static LDS (grids of up to MAXQ cells): qGH, the packed centres, vparGH
// (high edges, then GL0).  Dynamic LDS: the window of 16-bit phase-space
// counters (two per word, a.hwords words).  Loads and stores address the
// tube's segment through a uniform base and a 32-bit byte offset.
// (launch bounds: minimum resident wavefronts per SIMD, which caps the VGPRs)
#ifndef KAOS_FP64_WAVES_EU
#define KAOS_FP64_WAVES_EU 4
#endif
// CX: the instantiation launched when a tube has charge exchange; it stores
// cos(thetak4) of every ion of such a tube (the frame of a conversion at the
// next step), so the plain push carries no extra store.
// BLOCK: threads per block; each thread pushes KAOS_TILE / BLOCK ions of a
// tile (the autotuner times 256 and 512 for the plain instantiation).
template <int MAXQ, bool CX, int BLOCK = KAOS_FP64_BLOCK>
__global__ void __launch_bounds__(BLOCK, KAOS_FP64_WAVES_EU) k_push_fp64(StepArgs a) {
  constexpr int BLK = BLOCK, PPT = KAOS_TILE / BLOCK;
  __shared__ double s_qGH[MAXQ], s_vGH[KAOS_MAXV + 1];
  __shared__ KaosCE64 s_cE[MAXQ];
  extern __shared__ unsigned int s_hist[];
  __shared__ unsigned int s_pushed;
  __shared__ int s_wq0, s_wnq, s_wtube, s_group;
  __shared__ int4 s_wbox;
  const int tid = threadIdx.x;
  int cur = -1;
  const int sflags = (a.do_interp ? 1 : 0) | (a.do_count ? 2 : 0) | (a.do_remove ? 4 : 0) | (a.do_qk ? 8 : 0);
  // per-tube scalars and base pointers
  int hoff = 0, goff = 0, nq = 0, nvpar = 0, nvperp = 0, desc = 0;
  uint32_t tcount = 0;
  long long toff = 0;
  double *rsT = nullptr, *vpT = nullptr, *qkT = nullptr, *aeT = nullptr, *agT = nullptr;
  double* ck4T = nullptr;                  // CX instantiation: cos(thetak4) of a charge-exchange tube
  uint32_t* cellT = nullptr;
  double qGL0 = 0.0, qGHlast = 0.0, qsms = 0.0, vscale = 0.0;
  int eambsign = 0;
  uint32_t tbox = 0u;
  kaos_push_par_t P = kaos_push_par(1.0, 1.0, 1.0, 0.0, 0, 0, 0);
  for (int pass = 0;; ++pass) {
  const int g = kaos_next_group(a, pass, &s_group);
  if (g < 0) break;
  int t0 = g * a.tiles_per_block;
  int t1 = min(t0 + a.tiles_per_block, a.ntiles);
  unsigned int npushed = 0;
  if (tid == 0) {
    s_pushed = 0u;
    // histogram window: config cells [wq0, wq0 + wnq) x the velocity box of the
    // group's first tube (ions are stored near cell order and move less than a
    // cell per step)
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
  // software pipeline: the state of this thread's next ion (same tube) is
  // loaded while the current ion is pushed
  bool pf = false;
  uint32_t pf_j = 0u, pf_c = 0u;
  double pf_r = 0.0, pf_v = 0.0, pf_q = 0.0;
  for (int tile = t0; tile < t1; ++tile) {
    int tb = a.tile_tube[tile];
    if (tb != cur) {
      pf = false;
      __syncthreads();
      const DevTube& T = a.tubes[tb];
      nq = T.nq; nvpar = T.nvpar; nvperp = T.nvperp; desc = T.desc;
      goff = (int)T.goff; hoff = (int)T.hoff;
      qGL0 = T.qGL0; qGHlast = T.qGHlast; qsms = T.qsms; eambsign = T.eambsign;
      P = kaos_push_par(T.h, T.p, T.cphi, T.sphi, T.grav, T.eamb, T.eambsign);
      tbox = kaos_pack_box(a.vbox[tb]);
      for (int i = tid; i < nq; i += BLK) {
        s_qGH[i] = a.qGH[goff + i];
        KaosCE64 e;
        e.q = a.qGC[goff + i]; e.ea = a.EA[goff + i]; e.eg = a.EG[goff + i]; e.pad = 0.0;
        s_cE[i] = e;
      }
      for (int i = tid; i <= nvpar; i += BLK) s_vGH[i] = a.vparGH[T.voff + i];
      toff = T.off;
      tcount = (uint32_t)a.count[tb];
      rsT = a.rs + toff; vpT = a.vp + toff; qkT = a.qk + toff; aeT = a.ae + toff; agT = a.ag + toff;
      cellT = a.cell + toff;
      if (CX) ck4T = T.cx ? a.ck4 + toff : nullptr;
      vscale = (double)nvpar / (a.vparGH[T.voff + nvpar - 1] - a.vparGH[T.voff + nvpar]);
      cur = tb;
      __syncthreads();
    }
    const uint32_t start = (uint32_t)(a.tile_start[tile] - toff);
#pragma unroll 1
    for (int it = 0; it < PPT; ++it) {
      const uint32_t jl = start + (uint32_t)(it * BLK + tid);
      if (jl >= tcount) break;
      const uint32_t b8 = jl * 8u, b4 = jl * 4u;
      double rs, vs, wq;
      uint32_t cold;
      if (pf && pf_j == jl) { rs = pf_r; vs = pf_v; wq = pf_q; cold = pf_c; }
      else { rs = kx_at(rsT, b8); vs = kx_at(vpT, b8); wq = kx_at(qkT, b8); cold = kx_at(cellT, b4); }
      {
        uint32_t jn = 0u;
        bool nxt = false;
        if (it + 1 < PPT) { jn = jl + BLK; nxt = jn < tcount; }
        else if (tile + 1 < t1 && a.tile_tube[tile + 1] == cur) {
          jn = (uint32_t)(a.tile_start[tile + 1] - toff) + (uint32_t)tid;
          nxt = jn < tcount;
        }
        pf = nxt;
        if (nxt) {
          pf_j = jn;
          pf_r = kx_at(rsT, jn * 8u); pf_v = kx_at(vpT, jn * 8u); pf_q = kx_at(qkT, jn * 8u);
          pf_c = kx_at(cellT, jn * 4u);
        }
      }
      if (rs == 0.0) continue;
      const int qprev = (int)(cold >> 16) - 1;
      if ((sflags & 4) && (qprev == 0 || qprev == -1)) {
        // BoundaryConditionsSub at a master step: ions whose cell of the previous
        // push is 0 (LB) or -1 (UB) leave the arrays (BC:158-256, BCReIndex:33-133)
        kx_at(rsT, b8) = 0.0;
        atomicAdd(&a.esc_acc[2 * cur + (qprev == 0 ? 0 : 1)], 1ull);
        continue;
      }
      ++npushed;
      const int hint = (qprev >= 1) ? qprev : (qprev == 0 ? 1 : nq);
      double aea, ag;
      if (sflags & 1) {
        double ea, eg;
        if (kd_interp2_fast(wq, s_cE, nq, qGL0, qGHlast, desc, hint, &ea, &eg)) {
          aea = kaos_aea_of(ea, qsms, eambsign);
          ag = kaos_ag_of(eg, qsms);
          kx_at(aeT, b8) = aea; kx_at(agT, b8) = ag;
        } else {
          aea = kx_at(aeT, b8); ag = kx_at(agT, b8);
        }
      } else {
        aea = kx_at(aeT, b8); ag = kx_at(agT, b8);
      }
      kaos_line_t L = kaos_line_i(rs, P.irep);
      kaos_qb_t gA = kaos_stageA(L);
      const int qc = kd_find_qcell_fast(gA.q, s_qGH, qGL0, qGHlast, nq, desc, hint);
      kaos_push_out_t o;
      kaos_push_ion(L, gA, qc, vs, ag, aea, &P, &o);
      // velocity cells (2-D): the stored Vpar cell first, then the search;
      // off-grid keeps the previous index (KUA:1087-1127).  Vperp is constant on
      // this path (no wave heating), so its cell, found when the ion was binned
      // first (k_bin_initial, k_inject), is carried in the cell word.
      const int ipc = (int)(cold & 0xFFu), iq = (int)((cold >> 8) & 0xFFu);
      int ip = ipc;
      {
        const int ipx = max(ipc, 1);
        const double vlo = (ipx == 1) ? s_vGH[nvpar] : s_vGH[ipx - 2], vhi = s_vGH[ipx - 1];
        if (!(ipc >= 1 && ((ipc > 1) ? (o.vpar > vlo) : (o.vpar >= vlo)) && o.vpar <= vhi)) {
          ip = kaos_find_vcell_guess(o.vpar, s_vGH, s_vGH[nvpar], nvpar, vscale);
          if (ip == 0) ip = ipc;
        }
      }
      if ((sflags & 2) && qc >= 1) {
        if (ip >= 1 && iq >= 1) {
          const int wql = qc - wq0, jr = ip - wb.x, ir = iq - wb.z;
          if (cur == wtube && (unsigned)wql < (unsigned)wnq && (unsigned)jr < (unsigned)wb.y &&
              (unsigned)ir < (unsigned)wb.w) {
            const int li = (wql * wb.w + ir) * wb.y + jr;
            atomicAdd(&s_hist[li >> 1], 1u << (16 * (li & 1)));
          } else {
            atomicAdd(&a.hist[hoff + ((long long)(qc - 1) * nvperp + (iq - 1)) * nvpar + (ip - 1)], 1u);
            kaos_hbox_note(&a.hbox[goff + qc - 1], tbox, ip, iq);
          }
        } else {
          atomicAdd(&a.nqoff[goff + qc - 1], 1u);
        }
      }
      if (o.rsN != o.rsN || o.vparN != o.vparN) set_status(a.status, 1u << KAOS_E_NAN);
      kx_at(rsT, b8) = o.rsN;
      kx_at(vpT, b8) = o.vparN;
      if (sflags & 8) kx_at(qkT, b8) = o.qk4;
      if (CX && ck4T) kx_at(ck4T, b8) = o.zC / sqrt(o.u2C);
      kx_at(cellT, b4) = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)ip;
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

// ---------------------------------------------------------------------------
// ICR wave heating of a step in the FP64 form (WaveHeatingSub, called at
// stage A of the push, _8_1_1_KineticUpdateA.f90:1392), as a pass before the
// push over the pre-push state, in tubes with heating on (the mixed form
// heats inside its push, k_push_mixed<..., true>).  Per live ion: the Vperp cell of
// the cell word becomes the cell of the pre-kick |(Vperp1, Vperp2)| (the
// push bins with it, Vperpindk1, KUA:1108-1127; off the grid the previous
// index stays); an ion whose stage-A cell is inside the grid (WaveHeating:
// 32-33) gets the kick of kaos_icr_kick from the seeded Philox4x32 stream.
// The stage-A quantities (r, w, cell, V||) are the push's own: the same
// shared functions on the same state.
__global__ void __launch_bounds__(KAOS_BLOCK) k_icr(StepArgs a) {
  const int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  const int t = a.tile_tube[tile];
  const DevTube& T = a.tubes[t];
  if (!T.icr) return;
  const long long end = T.off + a.count[t];
  const long long step = *a.stepno;
  const double* GH = a.vperpGH + T.woff;
  const double irep = 1.0 / (KAOS_RE * T.p);
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    const long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    if (j >= end) break;
    const uint32_t cw = a.cell[j];
    const int qprev = (int)((cw >> 16) & KAOS_QMASK) - 1;
    const int hint = (qprev >= 1) ? qprev : (qprev == 0 ? 1 : T.nq);
    const double rs = a.rs[j];
    if (rs == 0.0) continue;
    const double vs = a.vp[j];
    const double v1 = a.vr[j], v2 = a.vr2[j];
    int iq = (int)((cw >> 8) & 0xFFu);
    const int iqn = kaos_find_vcell_from(fabs(sqrt(v1 * v1 + v2 * v2)), GH, GH[T.nvperp], T.nvperp, iq);
    if (iqn >= 1) iq = iqn;
    a.cell[j] = (cw & ~(0xFFu << 8)) | ((uint32_t)iq << 8);
    const kaos_line_t L = kaos_line_i(rs, irep);
    const kaos_qb_t gA = kaos_stageA(L);
    const int qc = kaos_find_qcell_from(gA.q, a.qGH + T.goff, T.qGL0, T.qGHlast, T.nq, T.desc, hint);
    if (qc == 0 || qc == -1) continue;
    const double vpar = gA.bR * (vs * gA.bR) + gA.bZ * (vs * gA.bZ);
    double u[4], v1N, v2N, vpN;
    kaos_icr_uniforms(a.seed, step, t, a.id[j], u);
    kaos_icr_kick(&T.icrp, L.r, L.w, vpar, v1, v2, u, &v1N, &v2N, &vpN);
    a.vr[j] = v1N;
    a.vr2[j] = v2N;
  }
}

// ---------------------------------------------------------------------------
// Initial step bookkeeping (KineticSolverASub, KSA:49-326): cells of the
// initial ions from the clamped q0 and the input Vpar0/Vperp0, counted into
// the master-step histogram.
__global__ void __launch_bounds__(KAOS_BLOCK) k_bin_initial(StepArgs a) {
  int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  int tb = a.tile_tube[tile];
  DevTube T = a.tubes[tb];
  long long end = T.off + a.count[tb];
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    if (j >= end) break;
    uint32_t cw = a.cell[j];
    int qc = (int)((cw >> 16) & KAOS_QMASK) - 1, iq = (int)((cw >> 8) & 0xFFu), ip = (int)(cw & 0xFFu);
    if (qc >= 1) {
      if (ip >= 1 && iq >= 1)
        atomicAdd(&a.hist[T.hoff + ((long long)(qc - 1) * T.nvperp + (iq - 1)) * T.nvpar + (ip - 1)], 1u);
      else
        atomicAdd(&a.nqoff[T.goff + qc - 1], 1u);
    }
  }
}

// Counts on the 3-D velocity grid (ION2VPERPflag = 1; IonParticleCountsSub's
// 3-D branch, _8_3_1_IonParticleCounts.f90:300-361) of a master step, as a
// pass over the pre-push state before the push and before the FP64 heating
// pass: the cells are taken from the pre-kick perpendicular components
// (KUA:985-1080, ahead of WaveHeatingSub at KUA:1392).  Per live ion the
// config cell and V|| are the push's stage-A ones (the same shared functions
// on the same binary64 state; the mixed form decodes its words first); at the
// first step (first = 1) the cells of the initial ions (KSA:49-326): the cell
// word's config and Vpar cells, found from the clamped q0 and Vpar0.  Ions
// the master push removes (previous cell 0 / -1) are not counted.  The three
// velocity cells are 0 off the grid (no stale index); an ion in a config
// cell with some velocity cell 0 counts in the config counts only.
//
// A block takes a group of KAOS_BIN3_TILES tiles (at most 64512 ions, so a
// 16-bit counter cannot overflow) and counts into an LDS window of whole
// config cells [wq0, wq0 + wnq) of the group's first tube, laid out as the
// global counts (so the flush adds a contiguous range); the window is centred
// on the cells of the group's first and last ions, which hold the group's
// ions when they were last sorted.  A thermal distribution puts most ions of
// a cell in a few hundred bins, so the window turns ~1 global atomic per ion
// into ~1 per occupied bin per block.  Ions outside the window (or of another
// tube) go to global atomics; no result depends on the window.  The cell
// searches run on LDS copies of the tube's edges, from the ion's previous
// config cell and from a uniform-spacing guess for the velocity cells: walks
// with the bisections' results (kaos_find_qcell_from, kaos_find_vcell_guess)
// and no dependent global loads.
#define KAOS_BIN3_BLOCK 1024
#define KAOS_BIN3_TILES 63
#ifndef KAOS_BIN3_UNROLL
#define KAOS_BIN3_UNROLL 4                     // tiles whose ions one thread loads together
#endif
#ifndef KAOS_BIN3_MAXWORDS
#define KAOS_BIN3_MAXWORDS 32768              // 128 KB of LDS: one 32x32x64 cell of 16-bit counters
#endif
static_assert(KAOS_TILE == KAOS_BIN3_BLOCK, "k_bin3d takes one ion of a tile per thread");
static_assert(KAOS_BIN3_TILES * KAOS_TILE < 65536, "16-bit window counters");
// The state words one ion of the binning pass reads (either form).
template <int MIXED> struct Bin3Ion;
template <> struct Bin3Ion<0> { uint32_t cw; double rs, vs, v1, v2; };
template <> struct Bin3Ion<1> { uint32_t cw; int rdi, vpi; float v1, v2; uint32_t c12, lo; };

__device__ __forceinline__ void bin3_load(const StepArgs& a, long long j, Bin3Ion<0>& x) {
  x.cw = a.cell[j]; x.rs = a.rs[j]; x.vs = a.vp[j]; x.v1 = a.vr[j]; x.v2 = a.vr2[j];
}
__device__ __forceinline__ void bin3_load(const StepArgs& a, long long j, Bin3Ion<1>& x) {
  x.cw = a.cell[j]; x.rdi = a.rdi[j]; x.vpi = a.vpi[j]; x.v1 = a.vrf[j]; x.v2 = a.vrf2[j]; x.lo = a.lo[j];
  x.c12 = a.c12[j];
}
__device__ __forceinline__ bool bin3_dead(const Bin3Ion<0>& x) { return x.rs == 0.0; }
__device__ __forceinline__ bool bin3_dead(const Bin3Ion<1>& x) { return x.cw == KAOS_DEAD; }

template <int MIXED>
__global__ void __launch_bounds__(KAOS_BIN3_BLOCK) k_bin3d(StepArgs a, int first, int wwords) {
  const int2 grp = a.grp3[blockIdx.x];
  extern __shared__ unsigned int s_w3[];
  __shared__ double s_qGH[KAOS_MAXQ], s_vGH[KAOS_MAXV + 1], s_v1GH[KAOS_MAXV + 1], s_v2GH[KAOS_MAXV + 1];
  __shared__ double s_rref[MIXED ? KAOS_MAXQ : 1];
  __shared__ int s_rk[MIXED ? KAOS_MAXQ : 1], s_vk[MIXED ? KAOS_MAXV + 1 : 1];
  __shared__ int s_wq0, s_wnq;
  const int t0 = grp.x, t1 = grp.y;
  if (t0 >= t1) return;                                   // uniform over the block
  const int tid = threadIdx.x;
  const int tb0 = a.tile_tube[t0];
  const DevTube& T0 = a.tubes[tb0];
  const long long nc0 = (long long)T0.nv1 * T0.nv2 * T0.nvpar;
  if (tid == 0) {
    const int wnq = (int)min((long long)T0.nq, (2LL * wwords) / nc0);
    s_wnq = wnq;
    s_wq0 = (wnq > 0) ? kaos_window_q0(a, t0, t1, tb0, T0.nq, T0.off + a.count[tb0], wnq) : 1;
  }
  for (int i = tid; i < wwords; i += KAOS_BIN3_BLOCK) s_w3[i] = 0u;
  __syncthreads();
  const int wq0 = s_wq0, wnq = s_wnq;
  int cur = -1;
  long long end = 0;
  double sp = 0.0, s1 = 0.0, s2 = 0.0;           // uniform-spacing guesses of the velocity grids
  // the tube's scalars in registers (the atomics below could alias the tube table)
  int nq = 0, desc = 0, nvpar = 0, nv1 = 0, nv2 = 0, heated = 0;
  long long goff = 0, h3off = 0, nc = 0;
  double irep = 0.0, qGL0 = 0.0, qGHlast = 0.0;
  // tiles in runs of up to KAOS_BIN3_UNROLL of one tube: every thread loads its
  // ion of each tile of the run before binning any (memory-level parallelism)
  for (int tile = t0; tile < t1;) {
    const int tb = a.tile_tube[tile];
    const DevTube& T = a.tubes[tb];
    if (tb != cur) {                               // uniform over the block: same tiles for every thread
      __syncthreads();
      for (int i = tid; i < T.nq; i += KAOS_BIN3_BLOCK) {
        s_qGH[i] = a.qGH[T.goff + i];
        if (MIXED) { s_rref[i] = a.rref[T.goff + i]; s_rk[i] = a.rk[T.goff + i]; }
      }
      for (int i = tid; i <= T.nvpar; i += KAOS_BIN3_BLOCK) {
        s_vGH[i] = a.vparGH[T.voff + i];
        if (MIXED) s_vk[i] = a.vk[T.voff + i];
      }
      for (int i = tid; i <= T.nv1; i += KAOS_BIN3_BLOCK) s_v1GH[i] = a.v1GH[T.v1off + i];
      for (int i = tid; i <= T.nv2; i += KAOS_BIN3_BLOCK) s_v2GH[i] = a.v2GH[T.v2off + i];
      const double* vGHg = a.vparGH + T.voff;
      sp = (double)T.nvpar / (vGHg[T.nvpar - 1] - vGHg[T.nvpar]);
      s1 = (double)T.nv1 / (a.v1GH[T.v1off + T.nv1 - 1] - a.v1GH[T.v1off + T.nv1]);
      s2 = (double)T.nv2 / (a.v2GH[T.v2off + T.nv2 - 1] - a.v2GH[T.v2off + T.nv2]);
      end = T.off + a.count[tb];
      nq = T.nq; desc = T.desc; nvpar = T.nvpar; nv1 = T.nv1; nv2 = T.nv2; heated = T.icr;
      goff = T.goff; h3off = T.h3off; nc = (long long)nv1 * nv2 * nvpar;
      irep = 1.0 / (KAOS_RE * T.p); qGL0 = T.qGL0; qGHlast = T.qGHlast;
      cur = tb;
      __syncthreads();
    }
    int k = 1;
    while (k < KAOS_BIN3_UNROLL && tile + k < t1 && a.tile_tube[tile + k] == tb) ++k;
    Bin3Ion<MIXED> xs[KAOS_BIN3_UNROLL];
    bool ok[KAOS_BIN3_UNROLL];
#pragma unroll
    for (int u = 0; u < KAOS_BIN3_UNROLL; ++u) {
      const long long j = a.tile_start[tile + min(u, k - 1)] + tid;   // KAOS_TILE == KAOS_BIN3_BLOCK
      ok[u] = (u < k) && (j < end);
      if (ok[u]) bin3_load(a, j, xs[u]);
    }
#pragma unroll
    for (int u = 0; u < KAOS_BIN3_UNROLL; ++u) {
      if (!ok[u]) continue;
      const Bin3Ion<MIXED>& x = xs[u];
      if (bin3_dead(x)) continue;
      const uint32_t cw = x.cw;
      int qc = (int)((cw >> 16) & KAOS_QMASK) - 1;
      int ip;
      if (first) {
        ip = (int)(cw & 0xFFu);
      } else {
        if (qc == 0 || qc == -1) continue;
        double rs, vs;
        if constexpr (MIXED) {
          const int Q = kx_refq(qc, nq);
          rs = s_rref[Q - 1] + kx_dec(x.rdi, x.lo, s_rk[Q - 1], (cw & KAOS_RFLAG) != 0);
          vs = kx_dec_v(x.vpi, x.lo >> 16, (int)(cw & 0xFFu), (cw & KAOS_VFLAG) != 0, s_vGH, s_vk, nvpar);
        } else {
          rs = x.rs;
          vs = x.vs;
        }
        const kaos_line_t L = kaos_line_i(rs, irep);       // kaos_line(rs, p), 1/(RE p) hoisted
        const kaos_qb_t gA = kaos_stageA(L);
        qc = kaos_find_qcell_from(gA.q, s_qGH, qGL0, qGHlast, nq, desc, qc);
        const double vpar = gA.bR * (vs * gA.bR) + gA.bZ * (vs * gA.bZ);
        ip = kaos_find_vcell_guess(vpar, s_vGH, s_vGH[nvpar], nvpar, sp);
      }
      if (qc < 1) continue;
      int i1 = 0, i2 = 0;
      bool have12 = false;
      if constexpr (MIXED) {
        if (!heated) {                               // cells of the binary64 components
          i1 = (int)(x.c12 & 0xFFu); i2 = (int)((x.c12 >> 8) & 0xFFu);
          have12 = true;
        }
      }
      if (!have12) {
        i1 = kaos_find_vcell_guess((double)x.v1, s_v1GH, s_v1GH[nv1], nv1, s1);
        i2 = kaos_find_vcell_guess((double)x.v2, s_v2GH, s_v2GH[nv2], nv2, s2);
      }
      if (ip >= 1 && i1 >= 1 && i2 >= 1) {
        const int r = ((i1 - 1) * nv2 + (i2 - 1)) * nvpar + (ip - 1);
        const int wql = qc - wq0;
        if (tb == tb0 && (unsigned)wql < (unsigned)wnq) {
          const int li = wql * (int)nc + r;                  // < 2 * KAOS_BIN3_MAXWORDS
          atomicAdd(&s_w3[li >> 1], 1u << (16 * (li & 1)));
        } else {
          atomicAdd(&a.hist3[h3off + (long long)(qc - 1) * nc + r], 1u);
        }
      } else {
        atomicAdd(&a.nqoff[goff + qc - 1], 1u);
      }
    }
    tile += k;
  }
  __syncthreads();
  if (wnq > 0) {
    unsigned int* gb = a.hist3 + T0.h3off + (long long)(wq0 - 1) * nc0;
    const long long nli = (long long)wnq * nc0;
    for (long long w = tid; 2 * w < nli; w += KAOS_BIN3_BLOCK) {
      const unsigned int v = s_w3[w];
      if (v & 0xFFFFu) atomicAdd(&gb[2 * w], v & 0xFFFFu);
      if (v >> 16) atomicAdd(&gb[2 * w + 1], v >> 16);      // zero beyond nli (never counted)
    }
  }
}

// Vperp1 | Vperp2 << 8 cells (0 off the grid) of a binary64 perpendicular
// velocity: the mixed form keeps them beside its binary32 components, so that
// its 3-D counts bin the components the FP64 form bins.
__device__ __forceinline__ uint16_t kd_cells12(const DevTube& T, const StepArgs& a, double v1, double v2) {
  const int i1 = kaos_find_vcell(v1, a.v1GH + T.v1off, a.v1GH[T.v1off + T.nv1], T.nv1);
  const int i2 = kaos_find_vcell(v2, a.v2GH + T.v2off, a.v2GH[T.v2off + T.nv2], T.nv2);
  return (uint16_t)(i1 | (i2 << 8));
}

// State of one initial ion in slot j: signed radius rs at dipole coordinate
// q, velocity (Vpar, Vperp1, Vperp2), |Vperp| = vperp.  The k0 cells
// (KSA:49-326) come from q clamped into the grid as KSA:55-129 does, and from
// Vpar and vperp.
__device__ __forceinline__ void kd_store_initial(const DevTube& T, const StepArgs& a, long long j, long long id, double q,
                                                 double rs, double vpar, double v1, double v2, double vperp,
                                                 const double* qGH, const double* vparGH, const double* vperpGH,
                                                 const double* rref, int mixed) {
  if (T.south) {
    if (q <= T.qGL0) q = T.qGL0;
    else if (q > T.qGHlast) q = T.qGHlast;
  } else {
    if (q > T.qGL0) q = T.qGL0;
    else if (q < T.qGHlast) q = T.qGHlast;
  }
  int qc = kaos_find_qcell(q, qGH, T.qGL0, T.qGHlast, T.nq, T.desc);
  a.id[j] = id;
  int ip = kaos_find_vcell(vpar, vparGH, vparGH[T.nvpar], T.nvpar);
  int iq = kaos_find_vcell(vperp, vperpGH, vperpGH[T.nvperp], T.nvperp);
  if (mixed) {
    int Q = kx_refq(qc, T.nq);
    bool rflag, vflag;
    uint32_t rlo, vlo;
    a.rdi[j] = kx_enc(rs - rref[Q - 1], a.rk[T.goff + Q - 1], rflag, rlo);
    a.vpi[j] = kx_enc_v(vpar, ip, vparGH, a.vk + T.voff, T.nvpar, vflag, vlo);
    a.lo[j] = rlo | (vlo << 16);
    a.vrf[j] = (float)v1;
    a.vrf2[j] = (float)v2;
    if (a.c12) a.c12[j] = kd_cells12(T, a, v1, v2);
    a.qkf[j] = (float)(q - kx_qref(a.qGC[T.goff + Q - 1]));
    a.aef[j] = 0.0f; a.agf[j] = 0.0f;
    a.cell[j] = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)ip |
                (rflag ? KAOS_RFLAG : 0u) | (vflag ? KAOS_VFLAG : 0u);
  } else {
    a.rs[j] = rs;
    a.vp[j] = vpar; a.vr[j] = v1; a.vr2[j] = v2; a.qk[j] = q;
    a.ae[j] = 0.0; a.ag[j] = 0.0;
    a.cell[j] = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)ip;
  }
}

// Upload: positions from the root solve of q0 (InitialConditionsSub),
// velocities Vpar0, Vperp1 = Vperp0, Vperp2 = 0.
__global__ void k_upload(DevTube T, long long n, const double* q0, const double* vpar0,
                         const double* vperp0, const double* qGH, const double* vparGH,
                         const double* vperpGH, const double* rref, StepArgs a, int mixed) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = q0[i];
  kd_store_initial(T, a, T.off + i, i, q, kaos_rs_of_q(q, T.p), vpar0[i], vperp0[i], 0.0, vperp0[i], qGH,
                   vparGH, vperpGH, rref, mixed);
}

// Per-ion state in the reference's arrays (kaos_upload_state), read the way
// KineticUpdateASub reads its Cartesian state: r = sqrt(x^2 + y^2 + z^2)
// (_8_1_1_KineticUpdateA.f90:117), the hemisphere from the sign of z, and
// Vpar = b-hat . V at that point (:978-980).  The cell word takes Qindk1 as
// given (the removal flag of the next master step and the search hint), the
// Vperp cell of |(Vperp1, Vperp2)| (the given one when that is off the grid:
// the 2-D branch's stale index, :1087-1127) and, in the binary64 form, the
// given Vparindk1 (the push finds the cell of the stored Vpar itself); the
// mixed form stores the cell of the stored Vpar beside its offset, as its push
// does.  Ions are numbered id0 + i.
struct UpState {
  const double *x, *y, *z, *vx, *vy, *vz, *v1, *v2, *vperp, *aea, *ag, *qk4;
  const long long *qind, *vparind, *vperpind;
};

__global__ void k_upload_state(DevTube T, long long n, UpState u, long long id0, const double* vparGH,
                               const double* vperpGH, const double* rref, StepArgs a, int mixed) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long j = T.off + i;
  const double x = u.x[i], y = u.y[i], z = u.z[i];
  const double r = sqrt((x * x + y * y) + z * z);
  const double rs = (z < 0.0 || (z == 0.0 && T.south)) ? -r : r;
  const kaos_line_t L = kaos_line(rs, T.p);
  const kaos_qb_t g = kaos_stageA(L);
  const double vpar = g.bR * (u.vx[i] * T.cphi + u.vy[i] * T.sphi) + g.bZ * u.vz[i];
  const double v1 = u.v1 ? u.v1[i] : u.vperp[i];
  const double v2 = u.v2 ? u.v2[i] : 0.0;
  const double vperp = u.v1 ? fabs(sqrt(v1 * v1 + v2 * v2)) : u.vperp[i];
  const int qc = (int)u.qind[i];
  int iq = kaos_find_vcell(vperp, vperpGH, vperpGH[T.nvperp], T.nvperp);
  if (iq < 1) iq = (int)u.vperpind[i];
  a.id[j] = id0 + i;
  if (mixed) {
    const int Q = kx_refq(qc, T.nq);
    bool rflag, vflag;
    uint32_t rlo, vlo;
    a.rdi[j] = kx_enc(rs - rref[Q - 1], a.rk[T.goff + Q - 1], rflag, rlo);
    int ip = kaos_find_vcell(vpar, vparGH, vparGH[T.nvpar], T.nvpar);
    a.vpi[j] = kx_enc_v(vpar, ip, vparGH, a.vk + T.voff, T.nvpar, vflag, vlo);
    a.lo[j] = rlo | (vlo << 16);
    if (ip < 1) ip = (int)u.vparind[i];
    a.vrf[j] = (float)v1;
    a.vrf2[j] = (float)v2;
    if (a.c12) a.c12[j] = kd_cells12(T, a, v1, v2);
    a.qkf[j] = (float)(u.qk4[i] - kx_qref(a.qGC[T.goff + Q - 1]));
    a.aef[j] = (float)u.aea[i]; a.agf[j] = (float)u.ag[i];
    a.cell[j] = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)ip |
                (rflag ? KAOS_RFLAG : 0u) | (vflag ? KAOS_VFLAG : 0u);
  } else {
    a.rs[j] = rs;
    a.vp[j] = vpar; a.vr[j] = v1; a.vr2[j] = v2; a.qk[j] = u.qk4[i];
    a.ae[j] = u.aea[i]; a.ag[j] = u.ag[i];
    a.cell[j] = ((uint32_t)(qc + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)u.vparind[i];
  }
}

// Initial ions generated on the device: ion i of the tube (cell by cell,
// coff = exclusive prefix of the per-cell counts, nq + 1 entries) sits at the
// centre of its cell (DensityProfileBSub) with a Maxwell-Boltzmann velocity of
// the cell's temperature (VelocityDistributionSub), from the seeded stream.
__global__ void k_seed(DevTube T, long long n, const long long* coff, const double* temp, unsigned long long seed,
                       int tube, const double* qGC, const double* qGH, const double* vparGH,
                       const double* vperpGH, const double* rref, StepArgs a, int mixed) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = T.nq - 1;                 // last cell Q with coff[Q] <= i
  while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (coff[mid] <= i) lo = mid; else hi = mid - 1; }
  const double q = qGC[lo];
  double u[6], rs, vpar, v1, v2, vperp;
  kaos_seed_uniforms(seed, tube, i, u);
  kaos_seed_ion(q, T.p, kaos_thermal_speed(temp[lo], T.ms), T.cphi, T.sphi, u, &rs, &vpar, &v1, &v2, &vperp);
  kd_store_initial(T, a, T.off + i, i, q, rs, vpar, v1, v2, vperp, qGH, vparGH, vperpGH, rref, mixed);
}

// Both perpendicular components of the ions just uploaded; the Vperp cell of
// the cell word follows from |(Vperp1, Vperp2)|.
__global__ void k_set_vperp(DevTube T, long long n, const double* v1, const double* v2, const double* vperpGH,
                            StepArgs a, int mixed) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long j = T.off + i;
  const double vp = fabs(sqrt(v1[i] * v1[i] + v2[i] * v2[i]));
  const int iq = kaos_find_vcell(vp, vperpGH, vperpGH[T.nvperp], T.nvperp);
  a.cell[j] = (a.cell[j] & ~(0xFFu << 8)) | ((uint32_t)iq << 8);
  if (mixed) {
    a.vrf[j] = (float)v1[i]; a.vrf2[j] = (float)v2[i];
    if (a.c12) a.c12[j] = kd_cells12(T, a, v1[i], v2[i]);
  } else {
    a.vr[j] = v1[i]; a.vr2[j] = v2[i];
  }
}

// Step counter on the device (the injection stream is keyed by the step), so
// that a captured step replays unchanged.
__global__ void k_step_tick(long long* stepno) { *stepno += 1; }

__global__ void k_fill_u32(unsigned int* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0u;
}

// Field of a master step (AmbipolarEfieldSub + GravfieldSub cell values), one
// block per tube; also clears the tube's escape counters, which the push of
// this master step then fills (BoundaryConditionsSub, BC:158-256).
// Start of a master step's bounding boxes of counts: every (tube, config
// cell) box is the tube's velocity box (the histogram itself is cleared whole
// by a streaming fill, which measured faster than clearing the boxes only).
__global__ void k_hbox_reset(int ncells, const DevTube* tubes, const int* blk_tube, const int* blk_q,
                             const int4* vbox, int4* hbox) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  const int t = blk_tube[ci];
  const int4 b = vbox[t];
  hbox[tubes[t].goff + blk_q[ci]] = make_int4(b.x, b.x + b.y - 1, b.z, b.z + b.w - 1);
}

// An input field (kaos_set_field) then replaces the computed one: at step 1
// (rotate_esc = 0) in both input modes, at later master steps in the fixed
// mode (_8_7_AmbipolarEfield.f90:1375-1407, after the stencils and the LB
// override).
__global__ void k_field(const DevTube* tubes, const double* hq, const double* dq, const double* Te,
                        const double* Mprev, double* M0F, double* M1F, double* EA, const double* EAin,
                        unsigned long long* esc_acc, unsigned long long* esc_rep, int rotate_esc) {
  int t = blockIdx.x;
  DevTube T = tubes[t];
  const double* M0 = Mprev + 6 * T.goff;  // M0 block of tube t
  double* nF = M0F + T.goff;
  double* E = EA + T.goff;
  for (int Q = threadIdx.x; Q < T.nq; Q += blockDim.x) {
    nF[Q] = kaos_moment_filter_cell(M0, T.nq, T.filter_pt, Q + 1);
    M1F[T.goff + Q] = kaos_moment_filter_cell(M0 + 2 * T.nq, T.nq, T.filter_pt_m1, Q + 1);
  }
  __syncthreads();
  for (int Q = threadIdx.x; Q < T.nq; Q += blockDim.x)
    E[Q] = kaos_ea_cell(Q, T.nq, T.south, T.eamb, T.eambsign, T.qs, hq + T.goff, dq + T.goff,
                        Te + T.goff, M0, nF);
  __syncthreads();
  if (threadIdx.x == 0) {
    kaos_ea_override(E, M0, T.eamb);
    if (rotate_esc) { esc_acc[2 * t] = 0ull; esc_acc[2 * t + 1] = 0ull; }
  }
  if (T.ea_in == KAOS_FIELD_INPUT_FIXED || (T.ea_in == KAOS_FIELD_INPUT_FIRST && !rotate_esc)) {
    __syncthreads();
    for (int Q = threadIdx.x; Q < T.nq; Q += blockDim.x) E[Q] = EAin[T.goff + Q];
  }
}

// LB/UB injection after the master push (KSB:195-793): ion k of tube t goes
// to slot off + count + k.  Uniforms from the caller's stream (6 per ion) or
// from the seeded Philox stream.
__global__ void k_inject(StepArgs a, const long long* inj_base, const long long* next_id, int ntubes,
                         const double* vperpGH, const double* uni,
                         const long long* uni_base, unsigned long long seed, const long long* stepno,
                         long long total, int mixed) {
  long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= total) return;
  int t = 0;
  while (t + 1 < ntubes && inj_base[t + 1] <= gi) ++t;
  DevTube T = a.tubes[t];
  long long k = gi - inj_base[t];
  long long nl = T.lbcond ? T.n_inj_lb : 0;
  int lb = (k < nl);
  int Qb = lb ? 0 : T.nq - 1;
  long long j = T.off + a.count[t] + k;
  double qc = a.qGC[T.goff + Qb];
  double u[6];
  if (uni) { for (int i = 0; i < 6; ++i) u[i] = uni[6 * (uni_base[t] + k) + i]; }
  else kaos_inject_uniforms(seed, *stepno, t, k, u);
  int qcell = kaos_find_qcell(qc, a.qGH + T.goff, T.qGL0, T.qGHlast, T.nq, T.desc);
  double r = kaos_rs_of_q(qc, T.p);
  kaos_line_t L = kaos_line(r, T.p);
  double vpar, v1, v2, vperp;
  kaos_mb_project(u, lb ? T.sig_lb : T.sig_ub, T.cphi, T.sphi, L, &vpar, &v1, &v2, &vperp);
  int ip = kaos_find_vcell(vpar, a.vparGH + T.voff, a.vparGH[T.voff + T.nvpar], T.nvpar);
  int iq = kaos_find_vcell(vperp, vperpGH + T.woff, vperpGH[T.woff + T.nvperp], T.nvperp);
  // a new slot's fields and stage-4 q are unset until its first push
  // (DataTypeReAllocA copies only the old ions, _8_0_6_DataTypeReAllocA.f90:
  // 259-313; the reference reads uninitialised memory there, 0 for most
  // batches): both fields start at 0 and q = 0 lies in no cell, so the next
  // master step's interpolation keeps them (_8_7_2_Gravfield.f90:305-420)
  // and the first push after injection runs without either field
  const double qk0 = 0.0;
  uint32_t flags = 0u;
  if (mixed) {
    int Q = kx_refq(qcell, T.nq);
    bool rflag, vflag;
    uint32_t rlo, vlo;
    a.rdi[j] = kx_enc(r - a.rref[T.goff + Q - 1], a.rk[T.goff + Q - 1], rflag, rlo);
    a.vpi[j] = kx_enc_v(vpar, ip, a.vparGH + T.voff, a.vk + T.voff, T.nvpar, vflag, vlo);
    a.lo[j] = rlo | (vlo << 16);
    a.vrf[j] = (float)v1;
    a.vrf2[j] = (float)v2;
    if (a.c12) a.c12[j] = kd_cells12(T, a, v1, v2);
    a.qkf[j] = (float)(qk0 - kx_qref(a.qGC[T.goff + Q - 1]));
    a.aef[j] = 0.0f; a.agf[j] = 0.0f;
    flags = (rflag ? KAOS_RFLAG : 0u) | (vflag ? KAOS_VFLAG : 0u);
  } else {
    a.rs[j] = r; a.vp[j] = vpar; a.vr[j] = v1; a.vr2[j] = v2; a.qk[j] = qk0;
    a.ae[j] = 0.0; a.ag[j] = 0.0;
  }
  a.cell[j] = ((uint32_t)(qcell + 1) << 16) | ((uint32_t)iq << 8) | (uint32_t)ip | flags;
  a.id[j] = next_id[t] + k;
  // charge exchange: a conversion before the ion's first push takes the frame
  // of its injection point (the reference leaves thetak4 of a new slot unset)
  if (a.ck4) a.ck4[j] = L.c;
  if (a.ck4f) a.ck4f[j] = (float)L.c;
  if (a.hist3) {
    // 3-D grid: the cells of the injected velocity (KSB:339-476), 0 off the grid
    if (qcell >= 1) {
      const int i1 = kaos_find_vcell(v1, a.v1GH + T.v1off, a.v1GH[T.v1off + T.nv1], T.nv1);
      const int i2 = kaos_find_vcell(v2, a.v2GH + T.v2off, a.v2GH[T.v2off + T.nv2], T.nv2);
      if (ip >= 1 && i1 >= 1 && i2 >= 1)
        atomicAdd(&a.hist3[T.h3off + (((long long)(qcell - 1) * T.nv1 + (i1 - 1)) * T.nv2 + (i2 - 1)) * T.nvpar +
                           (ip - 1)], 1u);
      else
        atomicAdd(&a.nqoff[T.goff + qcell - 1], 1u);
    }
    return;
  }
  if (qcell >= 1) {
    if (ip >= 1 && iq >= 1) {
      atomicAdd(&a.hist[T.hoff + ((long long)(qcell - 1) * T.nvperp + (iq - 1)) * T.nvpar + (ip - 1)], 1u);
      kaos_hbox_note(&a.hbox[T.goff + qcell - 1], kaos_pack_box(a.vbox[t]), ip, iq);
    } else {
      atomicAdd(&a.nqoff[T.goff + qcell - 1], 1u);
    }
  }
}

// Marginal counts of every tube over Vpar cells and over Vperp cells (inside
// the tube's bounding box of counts), one wavefront per (tube, config cell):
// lanes sum columns, then each row is summed across the lanes.
__global__ void __launch_bounds__(KAOS_BLOCK) k_marginals(const DevTube* tubes, const int* blk_tube,
    const int* blk_q, int ncells, const int4* hbox, const unsigned int* hist, unsigned int* vmp,
    unsigned int* vmr) {
  const int lane = threadIdx.x & 63;
  const int ci = blockIdx.x * (KAOS_BLOCK / 64) + (threadIdx.x >> 6);
  if (ci >= ncells) return;
  const int t = blk_tube[ci], Q = blk_q[ci];
  const DevTube& T = tubes[t];
  const int4 hb = hbox[T.goff + Q];
  if (hb.x > hb.y) return;
  const unsigned int* N = hist + T.hoff + (long long)Q * T.nvperp * T.nvpar;
  for (int b = hb.x - 1 + lane; b < hb.y; b += 64) {
    unsigned int x = 0u;
    for (int r = hb.z - 1; r < hb.w; ++r) x += N[r * T.nvpar + b];
    if (x) atomicAdd(&vmp[T.voff + b], x);
  }
  for (int r = hb.z - 1; r < hb.w; ++r) {
    unsigned int x = 0u;
    for (int b = hb.x - 1 + lane; b < hb.y; b += 64) x += N[r * T.nvpar + b];
    for (int st = 32; st >= 1; st >>= 1) x += __shfl_xor(x, st, 64);
    if (lane == 0 && x) atomicAdd(&vmr[T.woff + r], x);
  }
}

// Counts -> distribution function -> moments of one (tube, config cell) per
// block (_8_3_1, _8_4_1, _8_5, _8_6_1.._8_6_6), one pass over the counts: lane
// (g, b) sums column b over rows g, g + G, ...; lane b adds the G partials in
// g order and forms the column values; pairwise trees over the columns give
// the sums, in the oracle's order (kaos_mom_part / kaos_tree_sum).  Block
// (t, 0) also advances the tube's slot count and next id by this step's
// injected ions.
__global__ void __launch_bounds__(KAOS_BLOCK) k_moments(const DevTube* tubes, const int* blk_tube,
    const int* blk_q, const double* d3xC, const double* d3v, const double* dvpar, const double* dvperp,
    const double* vparGC, const double* vperpGC, const unsigned int* hist, const unsigned int* nqoff,
    unsigned int* nqc, double* NqRT, double* M, long long* count, long long* next_id,
    const long long* inj_cnt, const int* cells, const int4* hbox, long long* live, const unsigned long long* esc_acc) {
  __shared__ double s_P[KAOS_MOM_LANES][3];
  __shared__ unsigned int s_pc[KAOS_MOM_LANES];
  __shared__ double s_A[KAOS_MAXV + 1][3];
  __shared__ double s_x[6][KAOS_MAXV + 1];
  __shared__ unsigned int s_cc[KAOS_MAXV + 1];
  __shared__ double s_M[6];
  const int cell = cells[blockIdx.x];
  const int t = blk_tube[cell], Q = blk_q[cell];
  const DevTube& T = tubes[t];
  const int np = T.nvpar, G = kaos_mom_groups(np);
  const int tid = threadIdx.x;
  const double k = T.nsnormfac / d3xC[T.goff + Q];
  const uint32_t* N = hist + T.hoff + (long long)Q * T.nvperp * np;
  kaos_vgrid_t v = { np, T.nvperp, d3v + T.doff, dvperp + T.woff, dvpar + T.voff,
                     vperpGC + T.woff, vparGC + T.voff };
  const int4 hb = hbox[T.goff + Q];
  if (tid < G * np) {
    double P[3] = {0.0, 0.0, 0.0};
    uint32_t cs = 0u;
    const int b = tid % np;
    if (b + 1 >= hb.x && b + 1 <= hb.y) kaos_mom_part(&v, N, b, tid / np, G, hb.z - 1, hb.w - 1, k, P, &cs);
    s_P[tid][0] = P[0]; s_P[tid][1] = P[1]; s_P[tid][2] = P[2];
    s_pc[tid] = cs;
  }
  __syncthreads();
  if (tid < np) {
    double A[3] = {s_P[tid][0], s_P[tid][1], s_P[tid][2]};
    unsigned int cs = s_pc[tid];
    for (int g = 1; g < G; ++g) {
      for (int i = 0; i < 3; ++i) A[i] = A[i] + s_P[g * np + tid][i];
      cs += s_pc[g * np + tid];
    }
    for (int i = 0; i < 3; ++i) s_A[tid][i] = A[i];
    s_cc[tid] = cs;
    double o4[4];
    kaos_mom_col_first(v.vparGC[tid], A, o4);
    for (int i = 0; i < 4; ++i) s_x[i][tid] = o4[i];
  }
  __syncthreads();
  for (int st = 1; st < np; st *= 2) {
    if (tid < np && (tid % (2 * st)) == 0 && tid + st < np)
      for (int i = 0; i < 4; ++i) s_x[i][tid] = s_x[i][tid] + s_x[i][tid + st];
    __syncthreads();
  }
  if (tid == 0) {
    unsigned int tot = nqoff[T.goff + Q];
    for (int c = 0; c < np; ++c) tot += s_cc[c];
    nqc[T.goff + Q] = tot;
    NqRT[T.goff + Q] = (double)tot * k;
    double S[4] = {s_x[0][0], s_x[1][0], s_x[2][0], s_x[3][0]}, Mq[6] = {0, 0, 0, 0, 0, 0};
    kaos_mom_finish1(S, T.ms, Mq);
    for (int i = 0; i < 6; ++i) s_M[i] = Mq[i];
  }
  __syncthreads();
  if (tid < np) {
    double o2[2];
    kaos_mom_col_second(v.vparGC[tid], s_A[tid], s_M[2], o2);
    s_x[4][tid] = o2[0]; s_x[5][tid] = o2[1];
  }
  __syncthreads();
  for (int st = 1; st < np; st *= 2) {
    if (tid < np && (tid % (2 * st)) == 0 && tid + st < np) {
      s_x[4][tid] = s_x[4][tid] + s_x[4][tid + st];
      s_x[5][tid] = s_x[5][tid] + s_x[5][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double S2[2] = {s_x[4][0], s_x[5][0]}, Mq[6];
    for (int i = 0; i < 6; ++i) Mq[i] = s_M[i];
    kaos_mom_finish2(S2, T.ms, Mq);
    for (int i = 0; i < 6; ++i) M[6 * T.goff + (long long)i * T.nq + Q] = Mq[i];
    if (Q == 0 && inj_cnt) {
      count[t] += inj_cnt[t]; next_id[t] += inj_cnt[t];
      // live ions: + this step's injections - this push's removals
      live[t] += inj_cnt[t] - (long long)(esc_acc[2 * t] + esc_acc[2 * t + 1]);
    }
  }
}

// The 3-D moments (_8_6_1.._8_6_6, 3-D branches) of one (tube, config cell)
// per block, one pass over the counts [V1][V2][Vpar]: lane (g, b) sums column
// b over rows (V1, V2) g, g + G, ... (kaos_mom3_part); lane b adds the G
// partials in g order and forms the column values; pairwise trees over the
// columns give the sums in the oracle's order.  Writes the eight moments and
// the 2-D block's M0, M1Par, M2, M2Par (M1Perp, M2Perp: 0), which the next
// master step's fields read.  Block (t, 0) advances the tube's slot count,
// next id and live count as the 2-D kernels do.
__global__ void __launch_bounds__(KAOS_MOM3_LANES) k_moments3d(const DevTube* tubes, const int* blk_tube,
    const int* blk_q, const double* d3xC, const double* d3v3, const double* dvpar, const double* dv1,
    const double* dv2, const double* vparGC, const double* v1GC, const double* v2GC,
    const unsigned int* hist3, const unsigned int* nqoff, unsigned int* nqc, double* NqRT, double* M, double* M3,
    long long* count, long long* next_id, const long long* inj_cnt, long long* live,
    const unsigned long long* esc_acc) {
  __shared__ double s_P[KAOS_MOM3_LANES][5];
  __shared__ unsigned int s_pc[KAOS_MOM3_LANES];
  __shared__ double s_A[KAOS_MAXV + 1][5];
  __shared__ double s_x[8][KAOS_MAXV + 1];
  __shared__ unsigned int s_cc[KAOS_MAXV + 1];
  __shared__ double s_M[8];
  const int t = blk_tube[blockIdx.x], Q = blk_q[blockIdx.x];
  const DevTube& T = tubes[t];
  const int np = T.nvpar, G = kaos_mom3_groups(np);
  const int nrow = T.nv1 * T.nv2;
  const int tid = threadIdx.x;
  const double k = T.nsnormfac / d3xC[T.goff + Q];
  const uint32_t* N = hist3 + T.h3off + (long long)Q * nrow * np;
  kaos_vgrid3_t v = { np, T.nv1, T.nv2, d3v3 + T.d3off, dv1 + T.v1off, dv2 + T.v2off, dvpar + T.voff,
                      v1GC + T.v1off, v2GC + T.v2off, vparGC + T.voff };
  if (tid < G * np) {
    double P[5];
    uint32_t cs = 0u;
    kaos_mom3_part(&v, N, tid % np, tid / np, G, 0, nrow - 1, k, P, &cs);
    for (int i = 0; i < 5; ++i) s_P[tid][i] = P[i];
    s_pc[tid] = cs;
  }
  __syncthreads();
  if (tid < np) {
    double A[5];
    unsigned int cs = s_pc[tid];
    for (int i = 0; i < 5; ++i) A[i] = s_P[tid][i];
    for (int g = 1; g < G; ++g) {
      for (int i = 0; i < 5; ++i) A[i] = A[i] + s_P[g * np + tid][i];
      cs += s_pc[g * np + tid];
    }
    for (int i = 0; i < 5; ++i) s_A[tid][i] = A[i];
    s_cc[tid] = cs;
    double o6[6];
    kaos_mom3_col_first(v.vparGC[tid], A, o6);
    for (int i = 0; i < 6; ++i) s_x[i][tid] = o6[i];
  }
  __syncthreads();
  for (int st = 1; st < np; st *= 2) {
    if (tid < np && (tid % (2 * st)) == 0 && tid + st < np)
      for (int i = 0; i < 6; ++i) s_x[i][tid] = s_x[i][tid] + s_x[i][tid + st];
    __syncthreads();
  }
  if (tid == 0) {
    unsigned int tot = nqoff[T.goff + Q];
    for (int c = 0; c < np; ++c) tot += s_cc[c];
    nqc[T.goff + Q] = tot;
    NqRT[T.goff + Q] = (double)tot * k;
    double S[6], Mq[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 6; ++i) S[i] = s_x[i][0];
    kaos_mom3_finish1(S, T.ms, Mq);
    for (int i = 0; i < 8; ++i) s_M[i] = Mq[i];
  }
  __syncthreads();
  if (tid < np) {
    double A[5], o2[2];
    for (int i = 0; i < 5; ++i) A[i] = s_A[tid][i];
    kaos_mom3_col_second(v.vparGC[tid], A, s_M[3], o2);
    s_x[6][tid] = o2[0]; s_x[7][tid] = o2[1];
  }
  __syncthreads();
  for (int st = 1; st < np; st *= 2) {
    if (tid < np && (tid % (2 * st)) == 0 && tid + st < np) {
      s_x[6][tid] = s_x[6][tid] + s_x[6][tid + st];
      s_x[7][tid] = s_x[7][tid] + s_x[7][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double S2[2] = {s_x[6][0], s_x[7][0]}, Mq[8];
    for (int i = 0; i < 8; ++i) Mq[i] = s_M[i];
    kaos_mom3_finish2(S2, T.ms, Mq);
    for (int i = 0; i < 8; ++i) M3[8 * T.goff + (long long)i * T.nq + Q] = Mq[i];
    const double m6[6] = {Mq[0], 0.0, Mq[3], Mq[4], 0.0, Mq[7]};
    for (int i = 0; i < 6; ++i) M[6 * T.goff + (long long)i * T.nq + Q] = m6[i];
    if (Q == 0 && inj_cnt) {
      count[t] += inj_cnt[t]; next_id[t] += inj_cnt[t];
      live[t] += inj_cnt[t] - (long long)(esc_acc[2 * t] + esc_acc[2 * t + 1]);
    }
  }
}

// The same moments with one wavefront per config cell (grids of <= 64 Vpar
// cells, G = 1): lane b owns column b; the pairwise trees run across lanes
// with shuffles (lane i adds lane i + s for i = 0, 2s, 4s, ...), so the sums
// are kaos_tree_sum's.  `cells` lists the (tube, cell) blocks of such tubes.
__device__ __forceinline__ double kaos_wave_tree(double x, int lane, int n) {
  for (int st = 1; st < n; st *= 2) {
    double y = __shfl_down(x, st, 64);
    if ((lane % (2 * st)) == 0 && lane + st < n) x = x + y;
  }
  return x;
}

// Waves per SIMD of the moment kernel: the row loop waits on its count loads,
// so the kernel runs at full occupancy (62 VGPRs with the 4-row unroll).
#ifndef KAOS_MOM_WAVES_EU
#define KAOS_MOM_WAVES_EU 8
#endif
__global__ void __launch_bounds__(KAOS_BLOCK, KAOS_MOM_WAVES_EU) k_moments_wave(const DevTube* tubes, const int* blk_tube,
    const int* blk_q, const int* cells, int ncells, const double* d3xC, const double* d3v,
    const double* dvpar, const double* dvperp, const double* vparGC, const double* vperpGC,
    const unsigned int* hist, const unsigned int* nqoff, unsigned int* nqc, double* NqRT, double* M,
    long long* count, long long* next_id, const long long* inj_cnt, const int4* hbox, long long* live,
    const unsigned long long* esc_acc) {
  const int lane = threadIdx.x & 63;
  const int ci = blockIdx.x * (KAOS_BLOCK / 64) + (threadIdx.x >> 6);
  if (ci >= ncells) return;
  const int cell = cells[ci];
  const int t = blk_tube[cell], Q = blk_q[cell];
  const DevTube& T = tubes[t];
  const int np = T.nvpar;
  const double k = T.nsnormfac / d3xC[T.goff + Q];
  const uint32_t* N = hist + T.hoff + (long long)Q * T.nvperp * np;
  kaos_vgrid_t v = { np, T.nvperp, d3v + T.doff, dvperp + T.woff, dvpar + T.voff,
                     vperpGC + T.woff, vparGC + T.voff };
  double A[3] = {0.0, 0.0, 0.0}, o4[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t cs = 0;
  double uc = 0.0;
  const int4 hb = hbox[T.goff + Q];
  if (lane < np) {
    if (lane + 1 >= hb.x && lane + 1 <= hb.y) kaos_mom_part(&v, N, lane, 0, 1, hb.z - 1, hb.w - 1, k, A, &cs);
    uc = v.vparGC[lane];
    kaos_mom_col_first(uc, A, o4);
  }
  for (int st = 1; st < 64; st *= 2) cs += __shfl_xor(cs, st, 64);   // integer: any order
  double S[4];
  for (int i = 0; i < 4; ++i) S[i] = __shfl(kaos_wave_tree(o4[i], lane, np), 0, 64);
  double Mq[6] = {0, 0, 0, 0, 0, 0};
  kaos_mom_finish1(S, T.ms, Mq);
  double o2[2] = {0.0, 0.0};
  if (lane < np) kaos_mom_col_second(uc, A, Mq[2], o2);
  double S2[2];
  for (int i = 0; i < 2; ++i) S2[i] = __shfl(kaos_wave_tree(o2[i], lane, np), 0, 64);
  kaos_mom_finish2(S2, T.ms, Mq);
  if (lane == 0) {
    unsigned int tot = nqoff[T.goff + Q] + cs;
    nqc[T.goff + Q] = tot;
    NqRT[T.goff + Q] = (double)tot * k;
    for (int i = 0; i < 6; ++i) M[6 * T.goff + (long long)i * T.nq + Q] = Mq[i];
    if (Q == 0 && inj_cnt) {
      count[t] += inj_cnt[t]; next_id[t] += inj_cnt[t];
      // live ions: + this step's injections - this push's removals
      live[t] += inj_cnt[t] - (long long)(esc_acc[2 * t] + esc_acc[2 * t + 1]);
    }
  }
}

// Velocity box of each tube's LDS histogram window, from the marginal counts
// (k_marginals) of the master step just binned: the Vpar and Vperp index ranges that leave
// out at most 2.5e-5 of the binned ions at each end, widened by one cell for
// the next step's moves (all-zero marginals: the whole grid).  The marginals
// are cleared for the next master step.  Counts outside the box go to global
// atomics, so no result depends on it.
__device__ void kaos_box_range(const unsigned int* m, int n, int* lo, int* cnt) {
  unsigned long long tot = 0;
  for (int i = 0; i < n; ++i) tot += m[i];
  if (tot == 0) { *lo = 0; *cnt = n; return; }
  const unsigned long long thr = tot / 40000ull;
  unsigned long long c = 0;
  int a = 0;
  while (a < n - 1 && c + m[a] <= thr) { c += m[a]; ++a; }
  int b = n - 1;
  c = 0;
  while (b > a && c + m[b] <= thr) { c += m[b]; --b; }
  a = max(a - 1, 0);
  b = min(b + 1, n - 1);
  *lo = a; *cnt = b - a + 1;
}

__global__ void k_vbox(const DevTube* tubes, unsigned int* vmp, unsigned int* vmr, int4* box) {
  const DevTube& T = tubes[blockIdx.x];
  __shared__ int s_b[4];
  if (threadIdx.x == 0) {
    int lo, n;
    kaos_box_range(vmp + T.voff, T.nvpar, &lo, &n);
    s_b[0] = lo + 1; s_b[1] = n;
  } else if (threadIdx.x == 1) {
    int lo, n;
    kaos_box_range(vmr + T.woff, T.nvperp, &lo, &n);
    s_b[2] = lo + 1; s_b[3] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) box[blockIdx.x] = make_int4(s_b[0], s_b[1], s_b[2], s_b[3]);
  for (int i = threadIdx.x; i < T.nvpar; i += blockDim.x) vmp[T.voff + i] = 0u;
  for (int i = threadIdx.x; i < T.nvperp; i += blockDim.x) vmr[T.woff + i] = 0u;
}

// Reorder (compaction + counting sort by config cell), run every few master
// steps: live ions of each tube are packed to the front of its segment,
// grouped by their cell-word cell, so the push's block-local histogram window
// keeps covering a block's ions as they diffuse.  Dead slots are dropped (the
// physical form of BCReIndexSub's pack, _8_0_7_1_BCReIndex.f90:33-133); each
// ion keeps its id, which orders downloads as the reference orders its
// arrays.  Within a cell the new order depends on block scheduling; no result
// depends on the slot order (counts are integers, ions are independent).
__device__ __forceinline__ int reorder_key(const StepArgs& a, long long j, int mixed) {
  uint32_t cw = a.cell[j];
  if (mixed ? (cw == KAOS_DEAD) : (a.rs[j] == 0.0)) return -1;
  return (int)((cw >> 16) & KAOS_QMASK);     // qc + 1 in [0, nq + 1]
}

__global__ void __launch_bounds__(KAOS_BLOCK) k_reorder_count(StepArgs a, int mixed, unsigned int* keycnt) {
  __shared__ unsigned int s_cnt[KAOS_MAXQ + 2];
  int tile = blockIdx.x;
  int tb = a.tile_tube[tile];
  const DevTube& T = a.tubes[tb];
  int nk = T.nq + 2;
  for (int k = threadIdx.x; k < nk; k += KAOS_BLOCK) s_cnt[k] = 0u;
  __syncthreads();
  long long end = T.off + a.count[tb];
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    if (j >= end) break;
    int key = reorder_key(a, j, mixed);
    if (key >= 0) atomicAdd(&s_cnt[key], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nk; k += KAOS_BLOCK)
    if (s_cnt[k]) atomicAdd(&keycnt[T.koff + k], s_cnt[k]);
}

// One thread per tube: cursor = segment start + exclusive prefix of the key
// counts; the new ion count; the counts cleared for the next reorder.
__global__ void k_reorder_scan(const DevTube* tubes, int ntubes, unsigned int* keycnt,
                               unsigned long long* cursor, long long* cnt_new) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntubes) return;
  const DevTube& T = tubes[t];
  unsigned long long run = (unsigned long long)T.off;
  for (int k = 0; k < T.nq + 2; ++k) {
    cursor[T.koff + k] = run;
    run += keycnt[T.koff + k];
    keycnt[T.koff + k] = 0u;
  }
  cnt_new[t] = (long long)run - T.off;
}

__global__ void __launch_bounds__(KAOS_BLOCK) k_reorder_scatter(StepArgs a, StepArgs b, int mixed,
                                                                unsigned long long* cursor) {
  __shared__ unsigned int s_cnt[KAOS_MAXQ + 2];
  __shared__ unsigned long long s_base[KAOS_MAXQ + 2];
  int tile = blockIdx.x;
  int tb = a.tile_tube[tile];
  const DevTube& T = a.tubes[tb];
  int nk = T.nq + 2;
  for (int k = threadIdx.x; k < nk; k += KAOS_BLOCK) s_cnt[k] = 0u;
  __syncthreads();
  long long end = T.off + a.count[tb];
  int key[KAOS_TILE_PPT];
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    key[it] = (j < end) ? reorder_key(a, j, mixed) : -1;
    if (key[it] >= 0) atomicAdd(&s_cnt[key[it]], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nk; k += KAOS_BLOCK) {
    unsigned int n = s_cnt[k];
    s_base[k] = n ? atomicAdd(&cursor[T.koff + k], (unsigned long long)n) : 0ull;
    s_cnt[k] = 0u;
  }
  __syncthreads();
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    if (key[it] < 0) continue;
    long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    long long d = (long long)(s_base[key[it]] + atomicAdd(&s_cnt[key[it]], 1u));
    if (mixed) {
      b.rdi[d] = a.rdi[j]; b.vpi[d] = a.vpi[j]; b.vrf[d] = a.vrf[j]; b.vrf2[d] = a.vrf2[j]; b.lo[d] = a.lo[j];
      b.qkf[d] = a.qkf[j]; b.aef[d] = a.aef[j]; b.agf[d] = a.agf[j];
    } else {
      b.rs[d] = a.rs[j]; b.vp[d] = a.vp[j]; b.vr[d] = a.vr[j]; b.vr2[d] = a.vr2[j];
      b.qk[d] = a.qk[j]; b.ae[d] = a.ae[j]; b.ag[d] = a.ag[j];
    }
    b.cell[d] = a.cell[j];
    if (a.c12) b.c12[d] = a.c12[j];
    if (a.ck4) b.ck4[d] = a.ck4[j];
    if (a.ck4f) b.ck4f[d] = a.ck4f[j];
    b.id[d] = a.id[j];
  }
}

__global__ void k_reorder_finish(int ntubes, const long long* cnt_new, long long* count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntubes) count[t] = cnt_new[t];
}

// Cartesian state of a tube segment for download (zeros for a dead slot).
__global__ void k_reconstruct(DevTube T, long long n, StepArgs a, int mixed, double* out6) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long j = T.off + i;
  double r = 0.0, v = 0.0;
  if (mixed) {
    uint32_t cw = a.cell[j];
    if (cw != KAOS_DEAD) {
      int Q = kx_refq((int)((cw >> 16) & KAOS_QMASK) - 1, T.nq);
      r = a.rref[T.goff + Q - 1] + kx_dec(a.rdi[j], a.lo[j], a.rk[T.goff + Q - 1], (cw & KAOS_RFLAG) != 0);
      v = kx_dec_v(a.vpi[j], a.lo[j] >> 16, (int)(cw & 0xFFu), (cw & KAOS_VFLAG) != 0, a.vparGH + T.voff,
                   a.vk + T.voff,
                   T.nvpar);
    }
  } else {
    r = a.rs[j];
    v = a.vp[j];
  }
  double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0;
  if (r != 0.0) kaos_cartesian(r, v, T.p, T.cphi, T.sphi, &x, &y, &z, &vx, &vy, &vz);
  out6[6 * i + 0] = x; out6[6 * i + 1] = y; out6[6 * i + 2] = z;
  out6[6 * i + 3] = vx; out6[6 * i + 4] = vy; out6[6 * i + 5] = vz;
}

// ---------------------------------------------------------------------------
// Ion-neutral charge exchange (QEXCHANGEflagT = 1).
//
// Collision counts: k_cx_freq, one block per tube, at step 1 and every master
// step after the moments (IonNeutralCollisionFrequencyA/BSub through
// kaos_cx_cell).  Conversion, at the start of every later step: the first
// nint(poi) live ions of each cell in id order become ENAs (_8_KineticSolver.f90:
// 95-166).  The id threshold of each cell is found by a radix select over
// the ids in 14-bit digits, as many as the tube's ids need (ids below
// next_id: two digits up to 2^28 ids, three up to the 40-bit limit): per
// digit a histogram pass over the ions (k_cx_hist, counts gathered per
// wavefront and per block in LDS before the global atomics, since the ions
// of a tile share their cell and mostly their digit) and one wavefront per
// cell that finds the digit holding the remaining count (k_cx_scan); a cell
// with no more ions than its count takes all.  k_cx_convert then moves every ion at or below its cell's
// threshold to the tube's ENA slots (IonNeutralCollisionsSub) and frees its
// slot.  ENAs: k_ena_step removes those that left at LB/UB (at master
// steps), finds each one's cell at its pre-push position, counts them per
// cell at master steps and pushes them ballistically.

#define KAOS_ENA_DEAD ((int)0x80000000)
#define KAOS_CX_DBITS 14
#define KAOS_CX_BINS (1 << KAOS_CX_DBITS)
#define KAOS_CX_LEVELS 3                     // 42 bits >= the 40-bit id limit
#define KAOS_CX_SLOTS 64                     // per-block LDS table of (cell, digit) counts

// Digits of the radix select for ids below next_id (1..3).
__device__ __forceinline__ int kd_cx_levels(long long next_id) {
  const unsigned long long m = (unsigned long long)(next_id > 1 ? next_id - 1 : 1);
  const int bits = 64 - __clzll((long long)m);
  const int l = (bits + KAOS_CX_DBITS - 1) / KAOS_CX_DBITS;
  return l < 1 ? 1 : (l > KAOS_CX_LEVELS ? KAOS_CX_LEVELS : l);
}

// (signed radius, V||) of slot j of tube T in either form; false for a free slot.
__device__ __forceinline__ bool kd_line_state(const StepArgs& a, const DevTube& T, long long j, int mixed,
                                              double& r, double& v) {
  if (mixed) {
    const uint32_t cw = a.cell[j];
    if (cw == KAOS_DEAD) return false;
    const int Q = kx_refq((int)((cw >> 16) & KAOS_QMASK) - 1, T.nq);
    r = a.rref[T.goff + Q - 1] + kx_dec(a.rdi[j], a.lo[j], a.rk[T.goff + Q - 1], (cw & KAOS_RFLAG) != 0);
    v = kx_dec_v(a.vpi[j], a.lo[j] >> 16, (int)(cw & 0xFFu), (cw & KAOS_VFLAG) != 0, a.vparGH + T.voff,
                 a.vk + T.voff, T.nvpar);
    return true;
  }
  r = a.rs[j];
  v = a.vp[j];
  return r != 0.0;
}

__global__ void k_cx_freq(const DevTube* tubes, const double* M3, const double* NqRT, const unsigned int* nqc,
                          const double* d3x, const double* cxgeo, const double* nsneut, double* cx,
                          const long long* stepno, unsigned long long seed) {
  __shared__ double s_f[4][KAOS_MAXQ];
  const int t = blockIdx.x;
  const DevTube& T = tubes[t];
  if (!T.cx) return;
  const int nq = T.nq;
  const double* M = M3 + 8 * T.goff;
  const double *m1par = M + 3 * nq, *m1p1 = M + nq, *m1p2 = M + 2 * nq, *m2par = M + 7 * nq;
  if (T.cx_filter) {
    for (int Q = threadIdx.x; Q < nq; Q += blockDim.x) {
      s_f[0][Q] = kaos_moment_filter_cell(m1par, nq, T.filter_pt_m1, Q + 1);
      s_f[1][Q] = kaos_moment_filter_cell(m1p1, nq, T.cx_filter_pt, Q + 1);
      s_f[2][Q] = kaos_moment_filter_cell(m1p2, nq, T.cx_filter_pt, Q + 1);
      s_f[3][Q] = kaos_moment_filter_cell(m2par, nq, T.cx_filter_pt, Q + 1);
    }
    __syncthreads();
    m1par = s_f[0]; m1p1 = s_f[1]; m1p2 = s_f[2]; m2par = s_f[3];
  }
  const long long n = *stepno;
  for (int Q = threadIdx.x; Q < nq; Q += blockDim.x) {
    const long long g = T.goff + Q;
    kaos_cx_in_t I;
    I.m0 = M[Q]; I.m1par = m1par[Q]; I.m1p1 = m1p1[Q]; I.m1p2 = m1p2[Q]; I.m2par = m2par[Q];
    if (T.cx_filter && I.m0 == 0.0) I.m2par = 0.0;     // sigma is set only for populated cells
    I.cth = cxgeo[2 * g]; I.ell = cxgeo[2 * g + 1]; I.cphi = T.cphi; I.sphi = T.sphi;
    I.nsneut = nsneut[g]; I.h = T.h; I.sig_n = T.sig_n;
    I.NqRT = NqRT[g]; I.d3x = d3x[g]; I.nsnormfac = T.nsnormfac; I.nqc = (double)nqc[g];
    I.sig_scale = T.sig_scale;
    const kaos_cx_out_t r = kaos_cx_cell(&I, seed, n, t, Q + 1);
    cx[3 * T.goff + Q] = r.sigma;
    cx[3 * T.goff + nq + Q] = r.nu;
    cx[3 * T.goff + 2 * nq + Q] = r.poi;
  }
}

__global__ void k_cx_sel_init(int ncell, const DevTube* tubes, const int* blk_tube, const int* blk_q,
                              const double* cx, const long long* next_id, long long* need, long long* pref,
                              long long* thr, int* lev) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncell) return;
  const DevTube& T = tubes[blk_tube[ci]];
  const int Q = blk_q[ci];
  const long long g = T.goff + Q;
  need[g] = T.cx ? kaos_cx_take(cx[3 * T.goff + 2 * T.nq + Q]) : 0;
  pref[g] = 0;
  thr[g] = -1;
  lev[g] = kd_cx_levels(next_id[blk_tube[ci]]);
}

// The cell (Qindk1) of a live ion of a charge-exchange tube whose cell still
// selects, or 0.
__device__ __forceinline__ int kd_cx_cell(const StepArgs& a, const DevTube& T, long long j, int mixed,
                                          const long long* need) {
  const uint32_t cw = a.cell[j];
  if (mixed ? (cw == KAOS_DEAD) : (a.rs[j] == 0.0)) return 0;
  const int qc = (int)((cw >> 16) & KAOS_QMASK) - 1;
  if (qc < 1) return 0;
  return need[T.goff + qc - 1] > 0 ? qc : 0;
}

// Counts of (cell, digit) keys: each lane brings one key with a count (a run
// of its ions).  Up to KAOS_CX_ROUNDS times, the lanes of a wavefront that
// share the first remaining lane's key add their counts together into the
// block's LDS table (flushed to the global histogram at the end of the
// block); lanes left after that add theirs to the global histogram
// directly.  High digits are shared by whole tiles (the ions of a cell have
// neighbouring ids; ions that moved to a neighbouring cell bring a second or
// third key) and are gathered in the rounds; low digits are mostly distinct
// and go straight to distinct addresses.  The table starts filled with an
// invalid key, and a slot's key is written after it is reserved, so a reader
// sees either that invalid key or the right one: a reader that misses a key
// being added adds a second entry for it; a full table sends the count to
// the global histogram.
#define KAOS_CX_ROUNDS 4
__device__ __forceinline__ void kd_cx_count(unsigned int* hist, bool active, uint32_t key, unsigned int cnt,
                                            uint32_t* s_key, unsigned int* s_cnt, int* s_n) {
  const int lane = threadIdx.x & 63;
  for (int round = 0; round < KAOS_CX_ROUNDS; ++round) {
    const unsigned long long m = __ballot(active);
    if (m == 0ull) return;
    const int leader = __ffsll((long long)m) - 1;
    const uint32_t lk = (uint32_t)__shfl((int)key, leader, 64);
    const bool same = active && key == lk;
    unsigned int c = same ? cnt : 0u;
    for (int st = 32; st >= 1; st >>= 1) c += (unsigned int)__shfl_xor((int)c, st, 64);
    if (lane == leader) {
      const int n = min(*(volatile int*)s_n, KAOS_CX_SLOTS);
      int i = 0;
      while (i < n && ((volatile uint32_t*)s_key)[i] != lk) ++i;
      if (i < n) {
        atomicAdd(&s_cnt[i], c);
      } else {
        const int k = atomicAdd(s_n, 1);
        if (k < KAOS_CX_SLOTS) {
          atomicAdd(&s_cnt[k], c);
          ((volatile uint32_t*)s_key)[k] = lk;
        } else {
          atomicAdd(&hist[lk], c);
        }
      }
    }
    if (same) active = false;
  }
  if (active) atomicAdd(&hist[key], cnt);
}

// One digit of the id select: the tile's ions of selecting cells (loads for
// all of a lane's ions first), runs of equal keys within a lane merged, then
// counted per wavefront and block.  A tube whose ids need fewer digits than
// `level` + 1 returns at once.
__global__ void __launch_bounds__(KAOS_BLOCK) k_cx_hist(StepArgs a, int mixed, int level, const long long* need,
                                                        const long long* pref, const int* lev, unsigned int* hist) {
  __shared__ uint32_t s_key[KAOS_CX_SLOTS];
  __shared__ unsigned int s_cnt[KAOS_CX_SLOTS];
  __shared__ int s_n;
  const int tile = blockIdx.x, tb = a.tile_tube[tile];
  const DevTube& T = a.tubes[tb];
  if (!T.cx) return;
  const int L = lev[T.goff];                              // the same for every cell of the tube
  if (level >= L) return;
  if (threadIdx.x < KAOS_CX_SLOTS) { s_key[threadIdx.x] = 0xFFFFFFFFu; s_cnt[threadIdx.x] = 0u; }
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const long long end = T.off + a.count[tb];
  const int sh = KAOS_CX_DBITS * (L - 1 - level);
  bool on[KAOS_TILE_PPT];
  uint32_t key[KAOS_TILE_PPT];
#pragma unroll
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {
    const long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    on[it] = false;
    key[it] = 0u;
    if (j < end) {
      const int qc = kd_cx_cell(a, T, j, mixed, need);
      if (qc) {
        const long long g = T.goff + qc - 1;
        const long long id = a.id[j];
        if (level == 0 && (id >> 40) != 0) set_status(a.status, 1u << KAOS_E_CAPACITY);   // beyond 40-bit ids
        on[it] = level == 0 || (id >> (sh + KAOS_CX_DBITS)) == pref[g];
        key[it] = (uint32_t)(g * KAOS_CX_BINS + ((id >> sh) & (KAOS_CX_BINS - 1)));
      }
    }
  }
  unsigned int run[KAOS_TILE_PPT];
#pragma unroll
  for (int it = KAOS_TILE_PPT - 1; it >= 0; --it) {
    run[it] = on[it] ? 1u : 0u;
    if (it + 1 < KAOS_TILE_PPT && on[it] && on[it + 1] && key[it + 1] == key[it]) run[it] += run[it + 1];
  }
#pragma unroll
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {           // every lane runs every round (wave counts)
    const bool starts = on[it] && !(it > 0 && on[it - 1] && key[it - 1] == key[it]);
    kd_cx_count(hist, starts, key[it], run[it], s_key, s_cnt, &s_n);
  }
  __syncthreads();
  const int n = min(s_n, KAOS_CX_SLOTS);
  if ((int)threadIdx.x < n && s_cnt[threadIdx.x]) atomicAdd(&hist[s_key[threadIdx.x]], s_cnt[threadIdx.x]);
}

// One wavefront per cell: lane l holds digits 256 l .. 256 l + 255.
__global__ void k_cx_scan(int ncell, int level, const int* lev, long long* need, long long* pref, long long* thr,
                          unsigned int* hist) {
  constexpr int PER = KAOS_CX_BINS / 64;
  const int lane = threadIdx.x & 63;
  const long long g = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= ncell) return;
  const long long k = need[g];
  if (k <= 0 || level >= lev[g]) return;
  unsigned int* h = hist + g * KAOS_CX_BINS + PER * lane;
  unsigned long long sum = 0;
  for (int i = 0; i < PER; ++i) sum += h[i];
  unsigned long long incl = sum;
  for (int st = 1; st < 64; st <<= 1) {
    const unsigned long long y = __shfl_up(incl, st, 64);
    if (lane >= st) incl += y;
  }
  const unsigned long long total = __shfl(incl, 63, 64);
  const unsigned long long excl = incl - sum;
  const bool all = level == 0 && total <= (unsigned long long)k;     // every ion of the cell converts
  if (!all && excl < (unsigned long long)k && (unsigned long long)k <= incl) {   // exactly one lane
    unsigned long long c = excl;
    int d = 0;
    for (; d < PER; ++d) {
      if (c + h[d] >= (unsigned long long)k) break;
      c += h[d];
    }
    const long long np = (pref[g] << KAOS_CX_DBITS) | (long long)(PER * lane + d);
    pref[g] = np;
    if (level == lev[g] - 1) { thr[g] = np; need[g] = 0; }
    else need[g] = k - (long long)c;
  }
  for (int i = 0; i < PER; ++i) h[i] = 0u;                    // the histogram starts at zero next time
  if (all && lane == 0) { thr[g] = 0x7FFFFFFFFFFFFFFFLL; need[g] = 0; }
}

__global__ void __launch_bounds__(KAOS_BLOCK) k_cx_convert(StepArgs a, int mixed, const long long* thr, double* ena,
                                                           long long* eid, int* eq, unsigned int* evc,
                                                           long long* ena_cnt, long long ena_tot, long long* live) {
  const int tile = blockIdx.x, tb = a.tile_tube[tile];
  const DevTube& T = a.tubes[tb];
  if (!T.cx) return;
  const long long end = T.off + a.count[tb];
  const int lane = threadIdx.x & 63;
  for (int it = 0; it < KAOS_TILE_PPT; ++it) {           // every lane runs every iteration (wave slot counts)
    const long long j = a.tile_start[tile] + (long long)it * KAOS_BLOCK + threadIdx.x;
    bool conv = false;
    int qc = 0;
    if (j < end) {
      const uint32_t cw = a.cell[j];
      if (mixed ? (cw != KAOS_DEAD) : (a.rs[j] != 0.0)) {
        qc = (int)((cw >> 16) & KAOS_QMASK) - 1;
        if (qc >= 1) {
          const long long th = thr[T.goff + qc - 1];
          conv = th >= 0 && a.id[j] <= th;
        }
      }
    }
    // ENA slots of the wavefront's conversions: one atomic per wavefront
    const unsigned long long m = __ballot(conv);
    if (m == 0ull) continue;
    const int leader = __ffsll((long long)m) - 1;
    const int nconv = __popcll(m);
    long long base = 0;
    if (lane == leader) {
      base = (long long)atomicAdd((unsigned long long*)&ena_cnt[tb], (unsigned long long)nconv);
      const long long fit = min(max(T.ecap - base, 0LL), (long long)nconv);   // slots left for them
      if (fit > 0) atomicAdd((unsigned long long*)&live[tb], (unsigned long long)(-fit));
    }
    base = __shfl(base, leader, 64);
    if (!conv) continue;
    const long long e = base + __popcll(m & ((1ull << lane) - 1ull));
    if (e >= T.ecap) { set_status(a.status, 1u << KAOS_E_CAPACITY); continue; }
    double r, v;
    kd_line_state(a, T, j, mixed, r, v);                  // a converting ion is live
    const double v1 = mixed ? (double)a.vrf[j] : a.vr[j], v2 = mixed ? (double)a.vrf2[j] : a.vr2[j];
    const double c4 = mixed ? (double)a.ck4f[j] : a.ck4[j];
    double X[6];
    kaos_ena_from_ion(r, v, v1, v2, c4, T.p, T.cphi, T.sphi, X);
    const long long k = T.eoff + e;
    for (int i = 0; i < 6; ++i) ena[(long long)i * ena_tot + k] = X[i];
    eid[k] = a.id[j];
    eq[k] = qc;
    evc[k] = 0u;
    if (mixed) a.cell[j] = KAOS_DEAD;
    else a.rs[j] = 0.0;
  }
}

// The eight ENA moments of one (tube, config cell) per block at a master
// step: lane l of 256 sums the bins l, l + 256, ... (kaos_ena_mom_part), a
// pairwise tree over the lanes gives the sums (the oracle's order); the
// second moments take the first ones as their centre.
__global__ void __launch_bounds__(KAOS_ENA_LANES) k_ena_moments(const DevTube* tubes, const int* blk_tube,
    const int* blk_q, const unsigned int* nq_ena, const unsigned int* ena_hist, const double* evp,
    const double* evq, const double* evf, const double* hh, const double* d33v, const double* d3x, double* M) {
  __shared__ double s_x[4][KAOS_ENA_LANES];
  __shared__ double s_m1[3];
  const int ci = blockIdx.x;
  const DevTube& T = tubes[blk_tube[ci]];
  if (!T.cx) return;
  const int Q = blk_q[ci], nq = T.nq, tid = threadIdx.x;
  double* Mt = M + 8 * T.goff;
  if (nq_ena[T.goff + Q] == 0u) {              // NqENART = 0: F = 0 in every bin
    if (tid < 8) Mt[(long long)tid * nq + Q] = 0.0;
    return;
  }
  const int nb = T.nvp * T.nvq * T.nvphi;
  const double *GHp = evp + T.epoff, *GHq = evq + T.eqoff, *GHf = evf + T.efoff;
  kaos_enagrid_t g = {T.nvp, T.nvq, T.nvphi, GHp + 2 * (T.nvp + 1), GHq + 2 * (T.nvq + 1), GHf + 2 * (T.nvphi + 1),
                      GHp + (T.nvp + 1), GHq + (T.nvq + 1), GHf + (T.nvphi + 1), hh + T.eboff + (long long)Q * nb,
                      d33v + T.eboff + (long long)Q * nb};
  const uint32_t* cnt = ena_hist + T.ehoff + (long long)Q * nb;
  const double k = T.nsnormfac / d3x[T.goff + Q];
  double Mq[8];
  for (int stage = 1; stage <= 2; ++stage) {
    double m1[3] = {0.0, 0.0, 0.0}, o4[4];
    if (stage == 2) { m1[0] = s_m1[0]; m1[1] = s_m1[1]; m1[2] = s_m1[2]; }
    kaos_ena_mom_part(&g, cnt, tid, k, stage, m1, o4);
    for (int i = 0; i < 4; ++i) s_x[i][tid] = o4[i];
    __syncthreads();
    for (int st = 1; st < KAOS_ENA_LANES; st *= 2) {
      if ((tid % (2 * st)) == 0 && tid + st < KAOS_ENA_LANES)
        for (int i = 0; i < 4; ++i) s_x[i][tid] = s_x[i][tid] + s_x[i][tid + st];
      __syncthreads();
    }
    if (tid == 0) {
      const double S[4] = {s_x[0][0], s_x[1][0], s_x[2][0], s_x[3][0]};
      if (stage == 1) {
        kaos_ena_mom_finish1(S, Mq);
        s_m1[0] = Mq[1]; s_m1[1] = Mq[2]; s_m1[2] = Mq[3];
      } else {
        kaos_ena_mom_finish2(S, T.ms, Mq);
        for (int i = 0; i < 8; ++i) Mt[(long long)i * nq + Q] = Mq[i];
      }
    }
    __syncthreads();
  }
}

// Start of a master step's ENA outputs: counts per cell and escape counters.
__global__ void k_ena_reset(int ncell, int ntubes, unsigned int* nq_ena, unsigned long long* ena_esc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) nq_ena[i] = 0u;
  if (i < 2 * ntubes) ena_esc[i] = 0ull;
}

__global__ void k_ena_step(const DevTube* tubes, int ntubes, const long long* eoff, const long long* ena_cnt,
                           long long ena_tot, double* ena, int* eq, const double* qGH, unsigned int* nq_ena,
                           unsigned long long* ena_esc, int master, unsigned int* evc, const double* evp,
                           const double* evq, const double* evf, unsigned int* ena_hist) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ena_tot) return;
  int lo = 0, hi = ntubes;                   // the tube whose slots [eoff[t], eoff[t+1]) hold k
  while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (eoff[mid] <= k) lo = mid; else hi = mid; }
  const DevTube& T = tubes[lo];
  if (!T.cx) return;
  const long long e = k - T.eoff;
  if (e >= min(ena_cnt[lo], T.ecap)) return;
  const int q0 = eq[k];
  if (q0 == KAOS_ENA_DEAD) return;
  if (master && (q0 == 0 || q0 == -1)) {     // BC:309-470
    eq[k] = KAOS_ENA_DEAD;
    atomicAdd(&ena_esc[2 * lo + (q0 == 0 ? 0 : 1)], 1ull);
    return;
  }
  double X[6];
  for (int i = 0; i < 6; ++i) X[i] = ena[(long long)i * ena_tot + k];
  const double q = kaos_ena_q(X[0], X[1], X[2]);
  const int Q = (q == q) ? kaos_find_qcell(q, qGH + T.goff, T.qGL0, T.qGHlast, T.nq, T.desc) : 0;
  eq[k] = Q;
  if (Q >= 1) {
    // velocity cells of the pre-push state; a component off its grid keeps its cell
    double vp, vq, vf;
    kaos_ena_vcomp(X, &vp, &vq, &vf);
    const uint32_t old = evc[k];
    const double *GHp = evp + T.epoff, *GHq = evq + T.eqoff, *GHf = evf + T.efoff;
    int ip = kaos_find_vcell(vp, GHp, GHp[T.nvp], T.nvp);
    int iq = kaos_find_vcell(vq, GHq, GHq[T.nvq], T.nvq);
    int jf = kaos_find_vcell(vf, GHf, GHf[T.nvphi], T.nvphi);
    if (!ip) ip = (int)(old & 0xFFu);
    if (!iq) iq = (int)((old >> 8) & 0xFFu);
    if (!jf) jf = (int)((old >> 16) & 0xFFu);
    evc[k] = (uint32_t)ip | ((uint32_t)iq << 8) | ((uint32_t)jf << 16);
    if (master) {
      atomicAdd(&nq_ena[T.goff + Q - 1], 1u);
      if (ip && iq && jf)
        atomicAdd(&ena_hist[T.ehoff + (long long)(Q - 1) * T.nvp * T.nvq * T.nvphi +
                            ((long long)(ip - 1) * T.nvq + (iq - 1)) * T.nvphi + (jf - 1)], 1u);
    }
  }
  kaos_ena_push(X, T.h, T.grav);
  for (int i = 0; i < 6; ++i) ena[(long long)i * ena_tot + k] = X[i];
}

// ===========================================================================
// Host side.

struct HostTube {
  kaos_grid g;
  bool set = false, uploaded = false;
  long long cap = 0, count = 0;
  std::vector<double> qGL, qGH, qGC, hq, dq, d3x, rGC, cth, ell, Te;
  std::vector<double> vparGL, vparGH, vparGC, dvpar, vperpGL, vperpGH, vperpGC, dvperp, d3v;
  std::vector<double> v1GL, v1GH, v1GC, dv1, v2GL, v2GH, v2GC, dv2, d3v3;   // 3-D velocity grid
  std::vector<double> ea_in;                // input field (kaos_set_field), nq values
  int ea_mode = 0;                          // KAOS_FIELD_*
  std::vector<double> nsneut;               // charge exchange: nsnormCNeutGT per cell
  std::vector<double> vpGL, vpGH, vpGC, dvp, vqGL, vqGH, vqGC, dvq, vphiGL, vphiGH, vphiGC, dvphi, ena_hh, ena_d33v;
};

// One set of particle-state arrays (the binary64 or the mixed form's words).
struct StateBufs {
  double *rs = nullptr, *vp = nullptr, *vr = nullptr, *qk = nullptr, *ae = nullptr, *ag = nullptr;
  double* vr2 = nullptr;
  int *rdi = nullptr, *vpi = nullptr;
  uint32_t* lo = nullptr;
  float *vrf = nullptr, *qkf = nullptr, *aef = nullptr, *agf = nullptr;
  float* vrf2 = nullptr;
  uint32_t* cell = nullptr;
  uint16_t* c12 = nullptr;
  long long* id = nullptr;
  double* ck4 = nullptr;                     // charge exchange: cos(thetak4), binary64 form
  float* ck4f = nullptr;                     // the same, mixed form
};

struct kaos_ctx {
  kaos_cfg cfg;
  std::string err;
  std::vector<HostTube> ht;
  std::vector<DevTube> dt;
  bool built = false;
  long long last_step = 0;
  int last_master = 0;
  // device buffers
  DevTube* d_tubes = nullptr;
  int* d_tile_tube = nullptr; long long* d_tile_start = nullptr; int ntiles = 0;
  int* d_blk_tube = nullptr; int* d_blk_q = nullptr; int nblk_mom = 0;
  int *d_cells_wave = nullptr, *d_cells_block = nullptr; int n_cells_wave = 0, n_cells_block = 0;
  double *d_qGH = nullptr, *d_qGC = nullptr, *d_hq = nullptr, *d_dq = nullptr, *d_Te = nullptr,
         *d_d3x = nullptr, *d_EA = nullptr, *d_EAin = nullptr, *d_EG = nullptr, *d_M = nullptr, *d_M0F = nullptr,
         *d_M1F = nullptr, *d_NqRT = nullptr;
  double *d_vparGH = nullptr, *d_vperpGH = nullptr, *d_vparGC = nullptr, *d_vperpGC = nullptr,
         *d_dvpar = nullptr, *d_dvperp = nullptr, *d_d3v = nullptr;
  StateBufs st[2];                           // particle state; st[cur] is live, the other the reorder target
  int cur = 0;
  int *d_rk = nullptr, *d_vk = nullptr;
  unsigned int* d_keycnt = nullptr;          // reorder: ions per (tube, key)
  unsigned long long* d_cursor = nullptr;    // reorder: next free slot per (tube, key)
  long long *d_cnt_new = nullptr, *d_next_id = nullptr;
  long long nkeys_tot = 0;
  int masters_since_reorder = 0;
  double *d_rGH = nullptr, *d_rref = nullptr;
  int lds_nq = 0, lds_nv = 0;
  long long* d_count = nullptr;
  unsigned int *d_hist = nullptr, *d_nqc = nullptr, *d_nqoff = nullptr, *d_status = nullptr;
  unsigned long long *d_esc_acc = nullptr, *d_esc_rep = nullptr;
  long long *d_inj_base = nullptr, *d_inj_cnt = nullptr, *d_uni_base = nullptr;
  double* d_uni = nullptr; long long uni_cap = 0;
  long long ncell_tot = 0, nhist_tot = 0, cap_tot = 0;
  long long inj_total = 0;
  hipStream_t stream = nullptr;
  std::vector<long long> inj_base_h;
  unsigned long long* d_pushed = nullptr;
  // kernel timing (hipEvents on the library stream around each push launch)
  bool prof_on = false;
  // launch shape of the push kernels, the online autotuner and captured steps
  struct Launch { int tpb, hwords, persistent, block; };
  Launch launch = {KAOS_TPB, KAOS_HWORDS, 0, KAOS_PUSH_BLOCK_DEFAULT};
  unsigned int* d_work = nullptr;
  long long* d_live = nullptr;                  // live ions per tube (kaos_count)
  int any_icr = 0;                              // some tube has ICR heating on
  // charge exchange (some tube has kaos_grid.cx): per cell nsnormCNeutGT,
  // cos(thetaGCGT) and ellGCGT of the centre, sigma / nu / poi of the last
  // master step, the id selection's remaining count, prefix and threshold
  // and its 1024-bin digit histogram; the ENAs (SoA: x, y, z, Vx, Vy, Vz;
  // id; Qindk1, KAOS_ENA_DEAD for a free slot), their slots used per tube,
  // master-step counts per cell and LB/UB escapes per tube
  int any_cx = 0;
  long long ena_tot = 0;
  double *d_nsneut = nullptr, *d_cxgeo = nullptr, *d_cx = nullptr;
  long long *d_cx_need = nullptr, *d_cx_pref = nullptr, *d_cx_thr = nullptr;
  unsigned int* d_cx_hist = nullptr;
  int* d_cx_lev = nullptr;                      // radix-select digits of each cell's tube
  double* d_ena = nullptr;
  long long *d_eid = nullptr, *d_ena_cnt = nullptr, *d_eoff = nullptr;
  int* d_eq = nullptr;
  unsigned int* d_nq_ena = nullptr;
  unsigned long long* d_ena_esc = nullptr;
  // ENA velocity grid (per tube and axis [high edges, GL0 | centres | widths]
  // at DevTube.ep/eq/efoff; hh and d33v [Q][bin] at eboff), per-ENA packed
  // cells ip | iq << 8 | iphi << 16, phase-space counts [Q][bin] at ehoff and
  // the eight moments per cell ([8][nq] per tube at 8 goff)
  double *d_ena_vp = nullptr, *d_ena_vq = nullptr, *d_ena_vf = nullptr;
  double *d_ena_hh = nullptr, *d_ena_d33v = nullptr, *d_ena_M = nullptr;
  unsigned int *d_evc = nullptr, *d_ena_hist = nullptr;
  long long ena_hist_tot = 0;
  int4* d_vbox = nullptr;                       // per tube velocity box of the push's LDS window
  int4* d_hbox = nullptr;                       // per (tube, cell) bounding box of the master step's counts
  unsigned int *d_vmarg_par = nullptr, *d_vmarg_perp = nullptr;
  long long* d_stepno = nullptr;
  int n_cu = 256;
  bool tuning = false;
  size_t tune_i = 0;
  int tune_skip = 2;                            // master steps run before the first candidate
  std::vector<Launch> tune_cand;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tune_ev;
  std::map<unsigned long long, hipGraphExec_t> graphs;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_push, ev_step;
  // 3-D velocity grid (KAOS_FLAG_VPERP3D): Vperp1 / Vperp2 cells (high edges with
  // GL0 appended, centres, widths), d3v3, counts [Q][V1][V2][Vpar], 8 moments per cell
  bool v3d = false;
  double *d_v1GH = nullptr, *d_v1GC = nullptr, *d_dv1 = nullptr, *d_v2GH = nullptr, *d_v2GC = nullptr,
         *d_dv2 = nullptr, *d_d3v3 = nullptr, *d_M3 = nullptr;
  unsigned int* d_hist3 = nullptr;
  long long nhist3_tot = 0;
  int bin3_words = 0;                           // LDS window of k_bin3d (32-bit words)
  int nb3 = 0;                                  // binning blocks: groups of <= KAOS_BIN3_TILES tiles of one tube
  int2* d_grp3 = nullptr;
  // host mirrors sized at build time (the no-allocation step reads through them)
  std::vector<long long> cnt_h, live_h;
  bool in_noalloc = false;                      // inside kaos_step_noalloc: no tuning, no events
};

static std::string g_init_err;

// Scratch of kaos_step_noalloc: the caller's injection uniforms staged in a
// file-scope device array instead of the context's lazily allocated buffer.
// Every GPU holds its own copy for the life of the process; host code uses
// only the per-device address from hipGetSymbolAddress (the array's name in
// host code is not a device address).  All calls on one GPU share it, so
// kaos_step_noalloc calls that overlap on one GPU race.
#define KAOS_NOALLOC_UNIFORMS (6 * 65536)      // doubles: 6 per injected ion, 65536 ions per step (3 MiB)
#define KAOS_MAX_DEVICES 64
static __device__ __align__(256) double scratch[KAOS_NOALLOC_UNIFORMS];
static double* g_scratch[KAOS_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (once per device).
static hipError_t scratch_setup(int dev) {
  if (dev < 0 || dev >= KAOS_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[dev] = (double*)p;
  return hipSuccess;
}

#define HIPCHK(ctx, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);              \
      return KAOS_E_DEVICE;                                                        \
    }                                                                              \
  } while (0)

static int fail(kaos_ctx* c, int code, const std::string& m) { c->err = m; return code; }

extern "C" const char* kaos_last_error(kaos_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_init_err.c_str();
}

extern "C" int kaos_init(const kaos_cfg* cfg, kaos_ctx** out) {
  if (!cfg || !out) { g_init_err = "null argument"; return KAOS_E_ARG; }
  if (cfg->ntubes < 1 || cfg->ndatfac < 1 || (cfg->precision != KAOS_FP64 && cfg->precision != KAOS_MIXED)) {
    g_init_err = "bad kaos_cfg (ntubes >= 1, ndatfac >= 1, precision FP64|MIXED)";
    return KAOS_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    g_init_err = "no HIP device " + std::to_string(cfg->device);
    return KAOS_E_DEVICE;
  }
  if (hipSetDevice(cfg->device) != hipSuccess) { g_init_err = "hipSetDevice failed"; return KAOS_E_DEVICE; }
  if (scratch_setup(cfg->device) != hipSuccess) {
    g_init_err = "hipGetSymbolAddress of the no-allocation scratch failed";
    return KAOS_E_DEVICE;
  }
  kaos_ctx* c = new kaos_ctx();
  c->cfg = *cfg;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, cfg->device) == hipSuccess && ncu > 0)
    c->n_cu = ncu;
  if (!(cfg->flags & KAOS_FLAG_NO_AUTOTUNE)) {
    c->tuning = true;
    for (int pers = 0; pers < 2; ++pers)
      for (int hw : {4096, 6144})
        for (int tpb : {8, 16})
          for (int blk : {256, 512}) c->tune_cand.push_back({tpb, hw, pers, blk});
  }
  c->ht.resize(cfg->ntubes);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c; g_init_err = "hipStreamCreate failed"; return KAOS_E_DEVICE;
  }
  *out = c;
  return KAOS_OK;
}

static bool contiguous(const double* lo, const double* hi, int n) {
  for (int i = 0; i + 1 < n; ++i) if (!(lo[i + 1] == hi[i])) return false;
  return true;
}

extern "C" int kaos_set_grid(kaos_ctx* c, int32_t tube, const kaos_grid* g, int64_t capacity) {
  if (!c || !g) return KAOS_E_ARG;
  if (c->built) return fail(c, KAOS_E_STATE, "kaos_set_grid after the first upload");
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (g->nq < 11 || g->nq > KAOS_MAXQ) return fail(c, KAOS_E_ARG, "nq must be in [11, 512]");
  if (g->nvpar < 1 || g->nvpar > KAOS_MAXV || g->nvperp < 1 || g->nvperp > KAOS_MAXV)
    return fail(c, KAOS_E_ARG, "nvpar/nvperp must be in [1, 255]");
  if (g->filter_pt < 3 || (g->filter_pt % 2) == 0 || (g->filter_pt - 1) / 2 + 1 > g->nq / 2)
    return fail(c, KAOS_E_ARG, "filter_pt must be odd, >= 3 and fit the grid");
  if (g->filter_pt_m1par != 0 && (g->filter_pt_m1par < 3 || (g->filter_pt_m1par % 2) == 0 ||
                                  (g->filter_pt_m1par - 1) / 2 + 1 > g->nq / 2))
    return fail(c, KAOS_E_ARG, "filter_pt_m1par must be 0, or odd, >= 3 and fit the grid");
  if (capacity < 0 || capacity >= (1LL << 28))
    return fail(c, KAOS_E_ARG, "capacity must be in [0, 2^28) slots per tube");
  const double* need[] = {g->qGL, g->qGH, g->qGC, g->hqC, g->dqC, g->d3xC, g->rGC, g->cthGC, g->ellGC,
                          g->Te, g->vparGL, g->vparGH, g->vparGC, g->dvpar, g->vperpGL, g->vperpGH,
                          g->vperpGC, g->dvperp, g->d3v};
  for (const double* p : need) if (!p) return fail(c, KAOS_E_ARG, "null grid array");
  if (!contiguous(g->qGL, g->qGH, g->nq)) return fail(c, KAOS_E_ARG, "config edges not contiguous");
  if (!contiguous(g->vparGL, g->vparGH, g->nvpar) || !contiguous(g->vperpGL, g->vperpGH, g->nvperp))
    return fail(c, KAOS_E_ARG, "velocity edges not contiguous");
  int desc = g->qGL[0] > 0;
  for (int i = 0; i < g->nq; ++i) {
    bool ok = desc ? (g->qGL[i] > g->qGH[i] && g->qGL[i] >= g->qGC[i] && g->qGC[i] >= g->qGH[i])
                   : (g->qGL[i] < g->qGH[i] && g->qGL[i] <= g->qGC[i] && g->qGC[i] <= g->qGH[i]);
    if (!ok) return fail(c, KAOS_E_ARG, "config cell " + std::to_string(i + 1) + " not monotone");
  }
  if (g->icr) {
    const double pv[] = {g->icr_eta_lh, g->icr_xi_perp1, g->icr_xi_perp2, g->icr_s0, g->icr_f0};
    for (double v : pv)
      if (!(v > 0.0) || !std::isfinite(v)) return fail(c, KAOS_E_ARG, "ICR parameters must be positive");
    if (!(g->icr_lambda_perp >= 0.0) || !std::isfinite(g->icr_chi_perp1) || !std::isfinite(g->icr_chi_perp2))
      return fail(c, KAOS_E_ARG, "bad ICR wavelength or spectral index");
  }
  const bool v3d = (c->cfg.flags & KAOS_FLAG_VPERP3D) != 0;
  if (v3d) {
    if (g->nvperp1 < 1 || g->nvperp1 > KAOS_MAXV || g->nvperp2 < 1 || g->nvperp2 > KAOS_MAXV)
      return fail(c, KAOS_E_ARG, "nvperp1/nvperp2 must be in [1, 255]");
    const double* need3[] = {g->vperp1GL, g->vperp1GH, g->vperp1GC, g->dvperp1, g->vperp2GL, g->vperp2GH,
                             g->vperp2GC, g->dvperp2, g->d3v3};
    for (const double* p : need3) if (!p) return fail(c, KAOS_E_ARG, "null 3-D velocity grid array");
    if (!contiguous(g->vperp1GL, g->vperp1GH, g->nvperp1) || !contiguous(g->vperp2GL, g->vperp2GH, g->nvperp2))
      return fail(c, KAOS_E_ARG, "3-D velocity edges not contiguous");
    for (int i = 0; i < g->nvperp1; ++i)
      if (!(g->vperp1GL[i] < g->vperp1GH[i])) return fail(c, KAOS_E_ARG, "Vperp1 cells not increasing");
    for (int i = 0; i < g->nvperp2; ++i)
      if (!(g->vperp2GL[i] < g->vperp2GH[i])) return fail(c, KAOS_E_ARG, "Vperp2 cells not increasing");
  }
  if (g->cx) {
    if (!v3d)
      return fail(c, KAOS_E_ARG, "charge exchange needs KAOS_FLAG_VPERP3D (its ion mean velocity reads M1Perp1/M1Perp2)");
    if (!g->ns_neut) return fail(c, KAOS_E_ARG, "charge exchange needs ns_neut");
    if (!(g->t_neut > 0.0) || !(g->m_neut > 0.0) || !(g->cx_sigma_scale >= 0.0) || !std::isfinite(g->cx_sigma_scale))
      return fail(c, KAOS_E_ARG, "charge exchange needs t_neut > 0, m_neut > 0 and a finite cx_sigma_scale >= 0");
    if (g->cx_filter && (g->cx_filter_pt < 3 || (g->cx_filter_pt % 2) == 0 || (g->cx_filter_pt - 1) / 2 + 1 > g->nq / 2))
      return fail(c, KAOS_E_ARG, "cx_filter_pt must be odd, >= 3 and fit the grid");
    if (g->ena_cap < 0 || g->ena_cap >= (1LL << 28)) return fail(c, KAOS_E_ARG, "ena_cap must be in [0, 2^28)");
    for (int i = 0; i < g->nq; ++i)
      if (!(g->ns_neut[i] >= 0.0) || !std::isfinite(g->ns_neut[i]))
        return fail(c, KAOS_E_ARG, "ns_neut must be finite and >= 0");
    if (g->nvp < 1 || g->nvp > KAOS_MAXV || g->nvq < 1 || g->nvq > KAOS_MAXV || g->nvphi < 1 || g->nvphi > KAOS_MAXV)
      return fail(c, KAOS_E_ARG, "nvp/nvq/nvphi must be in [1, 255]");
    const double* ne[] = {g->vpGL, g->vpGH, g->vpGC, g->dvp, g->vqGL, g->vqGH, g->vqGC, g->dvq,
                          g->vphiGL, g->vphiGH, g->vphiGC, g->dvphi, g->ena_hh, g->ena_d33v};
    for (const double* p : ne) if (!p) return fail(c, KAOS_E_ARG, "null ENA velocity grid array");
    if (!contiguous(g->vpGL, g->vpGH, g->nvp) || !contiguous(g->vqGL, g->vqGH, g->nvq) ||
        !contiguous(g->vphiGL, g->vphiGH, g->nvphi))
      return fail(c, KAOS_E_ARG, "ENA velocity edges not contiguous");
  }
  HostTube& h = c->ht[tube];
  h.g = *g;
  int nq = g->nq, np = g->nvpar, nr = g->nvperp;
  if (g->cx) {
    h.nsneut.assign(g->ns_neut, g->ns_neut + nq);
    const int a = g->nvp, b = g->nvq, f = g->nvphi;
    const size_t nb = (size_t)a * b * f;
    h.vpGL.assign(g->vpGL, g->vpGL + a); h.vpGH.assign(g->vpGH, g->vpGH + a);
    h.vpGC.assign(g->vpGC, g->vpGC + a); h.dvp.assign(g->dvp, g->dvp + a);
    h.vqGL.assign(g->vqGL, g->vqGL + b); h.vqGH.assign(g->vqGH, g->vqGH + b);
    h.vqGC.assign(g->vqGC, g->vqGC + b); h.dvq.assign(g->dvq, g->dvq + b);
    h.vphiGL.assign(g->vphiGL, g->vphiGL + f); h.vphiGH.assign(g->vphiGH, g->vphiGH + f);
    h.vphiGC.assign(g->vphiGC, g->vphiGC + f); h.dvphi.assign(g->dvphi, g->dvphi + f);
    h.ena_hh.assign(g->ena_hh, g->ena_hh + nb * g->nq); h.ena_d33v.assign(g->ena_d33v, g->ena_d33v + nb * g->nq);
  } else {
    h.nsneut.clear();
  }
  if (v3d) {
    const int n1 = g->nvperp1, n2 = g->nvperp2;
    h.v1GL.assign(g->vperp1GL, g->vperp1GL + n1); h.v1GH.assign(g->vperp1GH, g->vperp1GH + n1);
    h.v1GC.assign(g->vperp1GC, g->vperp1GC + n1); h.dv1.assign(g->dvperp1, g->dvperp1 + n1);
    h.v2GL.assign(g->vperp2GL, g->vperp2GL + n2); h.v2GH.assign(g->vperp2GH, g->vperp2GH + n2);
    h.v2GC.assign(g->vperp2GC, g->vperp2GC + n2); h.dv2.assign(g->dvperp2, g->dvperp2 + n2);
    h.d3v3.assign(g->d3v3, g->d3v3 + (size_t)n1 * n2 * np);
  }
  h.qGL.assign(g->qGL, g->qGL + nq); h.qGH.assign(g->qGH, g->qGH + nq); h.qGC.assign(g->qGC, g->qGC + nq);
  h.hq.assign(g->hqC, g->hqC + nq); h.dq.assign(g->dqC, g->dqC + nq); h.d3x.assign(g->d3xC, g->d3xC + nq);
  h.rGC.assign(g->rGC, g->rGC + nq); h.cth.assign(g->cthGC, g->cthGC + nq); h.ell.assign(g->ellGC, g->ellGC + nq);
  h.Te.assign(g->Te, g->Te + nq);
  h.vparGL.assign(g->vparGL, g->vparGL + np); h.vparGH.assign(g->vparGH, g->vparGH + np);
  h.vparGC.assign(g->vparGC, g->vparGC + np); h.dvpar.assign(g->dvpar, g->dvpar + np);
  h.vperpGL.assign(g->vperpGL, g->vperpGL + nr); h.vperpGH.assign(g->vperpGH, g->vperpGH + nr);
  h.vperpGC.assign(g->vperpGC, g->vperpGC + nr); h.dvperp.assign(g->dvperp, g->dvperp + nr);
  h.d3v.assign(g->d3v, g->d3v + (size_t)np * nr);
  h.cap = capacity;
  h.set = true;
  return KAOS_OK;
}

// Allocate the device arrays once every tube's grid is known.
static bool plain_push(const kaos_ctx* c);

static int build_device(kaos_ctx* c) {
  int nt = c->cfg.ntubes;
  for (int t = 0; t < nt; ++t) if (!c->ht[t].set) return fail(c, KAOS_E_STATE, "tube " + std::to_string(t) + " has no grid");
  c->dt.resize(nt);
  long long goff = 0, voff = 0, woff = 0, hoff = 0, doff = 0, off = 0;
  std::vector<double> qGH, qGC, hq, dq, Te, d3x, EG, EAin, vparGHx, vperpGHx, vparGC, vperpGC, dvpar, dvperp, d3v;
  std::vector<double> rGH, rref;
  std::vector<int> rk, vk;
  std::vector<double> ena_vp, ena_vq, ena_vf, ena_hh, ena_d33;     // charge exchange: ENA grids
  std::vector<double> v1GHx, v1GC, dv1, v2GHx, v2GC, dv2, d3v3;
  long long v1off = 0, v2off = 0, h3off = 0, d3off = 0;
  const bool mixed = c->cfg.precision == KAOS_MIXED;
  std::vector<int> tile_tube, blk_tube, blk_q;
  std::vector<long long> tile_start;
  for (int t = 0; t < nt; ++t) {
    HostTube& h = c->ht[t];
    const kaos_grid& g = h.g;
    DevTube T;
    memset(&T, 0, sizeof(T));
    T.nq = g.nq; T.nvpar = g.nvpar; T.nvperp = g.nvperp; T.desc = g.qGL[0] > 0; T.south = !T.desc;
    T.grav = g.grav; T.eamb = g.eamb; T.eambsign = g.eambsign; T.filter_pt = g.filter_pt;
    T.filter_pt_m1 = g.filter_pt_m1par ? g.filter_pt_m1par : g.filter_pt;
    T.lbcond = g.lbcond; T.ubcond = g.ubcond;
    T.off = off; T.cap = h.cap; T.n_inj_lb = g.n_inj_lb; T.n_inj_ub = g.n_inj_ub;
    T.goff = goff; T.voff = voff; T.woff = woff; T.hoff = hoff; T.doff = doff;
    T.koff = goff + 2LL * t;
    T.p = g.p; T.cphi = g.cphi; T.sphi = g.sphi; T.h = g.h; T.ms = g.ms; T.qs = g.qs;
    T.qsms = g.qs / g.ms; T.nsnormfac = g.nsnormfac;
    T.sig_lb = sqrt((KAOS_KB * g.Ts_lb) / g.ms);
    T.sig_ub = sqrt((KAOS_KB * g.Ts_ub) / g.ms);
    T.qGL0 = g.qGL[0]; T.qGHlast = g.qGH[g.nq - 1];
    // radii of the q edges on the line (the mixed form bins |r| against them)
    T.rGL0 = fabs(kaos_rs_of_q(g.qGL[0], g.p));
    T.rGHlast = fabs(kaos_rs_of_q(g.qGH[g.nq - 1], g.p));
    T.rsgn = T.desc ? 1.0 : -1.0;
    T.icr = g.icr ? 1 : 0;
    if (T.icr) c->any_icr = 1;
    T.ea_in = h.ea_mode;
    if (g.cx) {
      T.cx = 1; T.cx_filter = g.cx_filter ? 1 : 0;
      T.cx_filter_pt = g.cx_filter_pt;
      T.sig_n = sqrt((KAOS_KB * g.t_neut) / g.m_neut);
      T.sig_scale = g.cx_sigma_scale;
      T.eoff = c->ena_tot; T.ecap = g.ena_cap;
      c->ena_tot += g.ena_cap;
      c->any_cx = 1;
      T.nvp = g.nvp; T.nvq = g.nvq; T.nvphi = g.nvphi;
      // per tube and axis a block [GH..., GL0 | GC..., 0 | widths..., 0] starting at e*off
      T.epoff = (long long)ena_vp.size(); T.eqoff = (long long)ena_vq.size(); T.efoff = (long long)ena_vf.size();
      auto axis = [](std::vector<double>& v, const std::vector<double>& GL, const std::vector<double>& GH,
                     const std::vector<double>& GC, const std::vector<double>& dv) {
        const size_t n = GH.size(), base = v.size();
        v.resize(base + 3 * (n + 1), 0.0);
        for (size_t i = 0; i < n; ++i) { v[base + i] = GH[i]; v[base + (n + 1) + i] = GC[i]; v[base + 2 * (n + 1) + i] = dv[i]; }
        v[base + n] = GL[0];
      };
      axis(ena_vp, h.vpGL, h.vpGH, h.vpGC, h.dvp);
      axis(ena_vq, h.vqGL, h.vqGH, h.vqGC, h.dvq);
      axis(ena_vf, h.vphiGL, h.vphiGH, h.vphiGC, h.dvphi);
      T.eboff = (long long)ena_hh.size();
      ena_hh.insert(ena_hh.end(), h.ena_hh.begin(), h.ena_hh.end());
      ena_d33.insert(ena_d33.end(), h.ena_d33v.begin(), h.ena_d33v.end());
      T.ehoff = c->ena_hist_tot;
      c->ena_hist_tot += (long long)g.nq * g.nvp * g.nvq * g.nvphi;
    }
    T.icrp.qs = g.qs; T.icrp.ms = g.ms; T.icrp.h = g.h;
    T.icrp.lambda_perp = g.icr_lambda_perp; T.icrp.eta_lh = g.icr_eta_lh;
    T.icrp.xi1 = g.icr_xi_perp1; T.icrp.xi2 = g.icr_xi_perp2; T.icrp.s0 = g.icr_s0;
    T.icrp.omega_g0 = KAOS_TWOPI * g.icr_f0;
    T.icrp.chi1 = g.icr_chi_perp1; T.icrp.chi2 = g.icr_chi_perp2;
    {
      double sds = 0.0;       // heating region: the whole tube, sum of hqC dqC in cell order
      for (int i = 0; i < g.nq; ++i) sds = sds + h.hq[i] * h.dq[i];
      T.icrp.sum_ds = sds;
    }
    // fixed-point exponent of a cell of width wd: 2^k wd < 2^30, so offsets up to
    // two widths fit in 31 bits
    auto fx = [](double wd) { int e = 0; frexp(wd, &e); return 30 - e; };
    double rprev = T.rGL0;
    for (int i = 0; i < g.nq; ++i) {
      double re = fabs(kaos_rs_of_q(h.qGH[i], g.p));
      if (mixed && !(re > rprev))
        return fail(c, KAOS_E_ARG, "mixed form: tube " + std::to_string(t) +
                                     " needs radii increasing with the cell index");
      rGH.push_back(re);
      rref.push_back(kaos_rs_of_q(h.qGC[i], g.p));
      rk.push_back(fx(re - rprev));
      rprev = re;
    }
    for (int i = 0; i < g.nvpar; ++i) vk.push_back(fx(h.vparGH[i] - h.vparGL[i]));
    vk.push_back(0);
    c->lds_nq = std::max(c->lds_nq, g.nq);
    c->lds_nv = std::max(c->lds_nv, g.nvpar);
    if (c->cfg.flags & KAOS_FLAG_VPERP3D) {
      T.nv1 = g.nvperp1; T.nv2 = g.nvperp2;
      T.v1off = v1off; T.v2off = v2off; T.h3off = h3off; T.d3off = d3off;
      for (int i = 0; i < T.nv1; ++i) { v1GHx.push_back(h.v1GH[i]); v1GC.push_back(h.v1GC[i]); dv1.push_back(h.dv1[i]); }
      v1GHx.push_back(h.v1GL[0]); v1GC.push_back(0.0); dv1.push_back(0.0);
      for (int i = 0; i < T.nv2; ++i) { v2GHx.push_back(h.v2GH[i]); v2GC.push_back(h.v2GC[i]); dv2.push_back(h.dv2[i]); }
      v2GHx.push_back(h.v2GL[0]); v2GC.push_back(0.0); dv2.push_back(0.0);
      d3v3.insert(d3v3.end(), h.d3v3.begin(), h.d3v3.end());
      v1off += T.nv1 + 1; v2off += T.nv2 + 1;
      d3off += (long long)T.nv1 * T.nv2 * g.nvpar;
      h3off += (long long)g.nq * T.nv1 * T.nv2 * g.nvpar;
    }
    c->dt[t] = T;
    for (int i = 0; i < g.nq; ++i) {
      qGH.push_back(h.qGH[i]); qGC.push_back(h.qGC[i]); hq.push_back(h.hq[i]); dq.push_back(h.dq[i]);
      Te.push_back(h.Te[i]); d3x.push_back(h.d3x[i]);
      EG.push_back(kaos_grav_cell(g.grav, g.ms, g.qs, h.cth[i], h.rGC[i], h.ell[i]));
      EAin.push_back(h.ea_in.empty() ? 0.0 : h.ea_in[i]);
      blk_tube.push_back(t); blk_q.push_back(i);
    }
    // velocity high edges followed by the low edge of cell 1 (read as GL0)
    for (int i = 0; i < g.nvpar; ++i) { vparGHx.push_back(h.vparGH[i]); vparGC.push_back(h.vparGC[i]); dvpar.push_back(h.dvpar[i]); }
    vparGHx.push_back(h.vparGL[0]); vparGC.push_back(0.0); dvpar.push_back(0.0);
    for (int i = 0; i < g.nvperp; ++i) { vperpGHx.push_back(h.vperpGH[i]); vperpGC.push_back(h.vperpGC[i]); dvperp.push_back(h.dvperp[i]); }
    vperpGHx.push_back(h.vperpGL[0]); vperpGC.push_back(0.0); dvperp.push_back(0.0);
    for (size_t i = 0; i < h.d3v.size(); ++i) d3v.push_back(h.d3v[i]);
    for (long long s = 0; s < h.cap; s += KAOS_TILE) { tile_tube.push_back(t); tile_start.push_back(off + s); }
    goff += g.nq; voff += g.nvpar + 1; woff += g.nvperp + 1; doff += (long long)g.nvpar * g.nvperp;
    hoff += (long long)g.nq * g.nvpar * g.nvperp; off += h.cap;
  }
  c->ncell_tot = goff; c->nhist_tot = hoff; c->cap_tot = off; c->ntiles = (int)tile_tube.size();
  c->nblk_mom = (int)blk_tube.size();
  // moment blocks: a wavefront per cell where the Vpar grid fits one, else a block
  std::vector<int> cw, cb;
  for (int i = 0; i < c->nblk_mom; ++i) (c->dt[blk_tube[i]].nvpar <= 64 ? cw : cb).push_back(i);
  c->n_cells_wave = (int)cw.size(); c->n_cells_block = (int)cb.size();
  auto up = [&](void** d, const void* h, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(d, bytes ? bytes : 8);
    if (e != hipSuccess) return e;
    if (h && bytes) e = hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
    return e;
  };
  HIPCHK(c, up((void**)&c->d_tubes, c->dt.data(), sizeof(DevTube) * nt));
  HIPCHK(c, up((void**)&c->d_tile_tube, tile_tube.data(), sizeof(int) * tile_tube.size()));
  HIPCHK(c, up((void**)&c->d_tile_start, tile_start.data(), sizeof(long long) * tile_start.size()));
  HIPCHK(c, up((void**)&c->d_blk_tube, blk_tube.data(), sizeof(int) * blk_tube.size()));
  HIPCHK(c, up((void**)&c->d_blk_q, blk_q.data(), sizeof(int) * blk_q.size()));
  HIPCHK(c, up((void**)&c->d_cells_wave, cw.data(), sizeof(int) * cw.size()));
  HIPCHK(c, up((void**)&c->d_cells_block, cb.data(), sizeof(int) * cb.size()));
  HIPCHK(c, up((void**)&c->d_qGH, qGH.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_qGC, qGC.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_hq, hq.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_dq, dq.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_Te, Te.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_d3x, d3x.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_EG, EG.data(), sizeof(double) * goff));
  std::vector<double> zeros(6 * goff, 0.0);
  HIPCHK(c, up((void**)&c->d_EA, zeros.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_EAin, EAin.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_M0F, zeros.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_M1F, zeros.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_NqRT, zeros.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_M, zeros.data(), sizeof(double) * 6 * goff));
  HIPCHK(c, up((void**)&c->d_vparGH, vparGHx.data(), sizeof(double) * voff));
  HIPCHK(c, up((void**)&c->d_vparGC, vparGC.data(), sizeof(double) * voff));
  HIPCHK(c, up((void**)&c->d_dvpar, dvpar.data(), sizeof(double) * voff));
  HIPCHK(c, up((void**)&c->d_vperpGH, vperpGHx.data(), sizeof(double) * woff));
  HIPCHK(c, up((void**)&c->d_vperpGC, vperpGC.data(), sizeof(double) * woff));
  HIPCHK(c, up((void**)&c->d_dvperp, dvperp.data(), sizeof(double) * woff));
  HIPCHK(c, up((void**)&c->d_d3v, d3v.data(), sizeof(double) * doff));
  HIPCHK(c, up((void**)&c->d_rGH, rGH.data(), sizeof(double) * goff));
  HIPCHK(c, up((void**)&c->d_rref, rref.data(), sizeof(double) * goff));
  size_t nslot = (size_t)(off ? off : 1);
  HIPCHK(c, up((void**)&c->d_rk, rk.data(), sizeof(int) * goff));
  HIPCHK(c, up((void**)&c->d_vk, vk.data(), sizeof(int) * voff));
  for (int b = 0; b < 2; ++b) {
    StateBufs& B = c->st[b];
    if (mixed) {
      HIPCHK(c, hipMalloc((void**)&B.rdi, sizeof(int) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vpi, sizeof(int) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.lo, sizeof(uint32_t) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vrf, sizeof(float) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vrf2, sizeof(float) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.qkf, sizeof(float) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.aef, sizeof(float) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.agf, sizeof(float) * nslot));
    } else {
      HIPCHK(c, hipMalloc((void**)&B.rs, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vp, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vr, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.vr2, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.qk, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.ae, sizeof(double) * nslot));
      HIPCHK(c, hipMalloc((void**)&B.ag, sizeof(double) * nslot));
      HIPCHK(c, hipMemset(B.rs, 0, sizeof(double) * nslot));
    }
    HIPCHK(c, hipMalloc((void**)&B.cell, sizeof(uint32_t) * nslot));
    HIPCHK(c, hipMemset(B.cell, 0xFF, sizeof(uint32_t) * nslot));
    HIPCHK(c, hipMalloc((void**)&B.id, sizeof(long long) * nslot));
    if (mixed && (c->cfg.flags & KAOS_FLAG_VPERP3D)) {
      HIPCHK(c, hipMalloc((void**)&B.c12, sizeof(uint16_t) * nslot));
      HIPCHK(c, hipMemset(B.c12, 0, sizeof(uint16_t) * nslot));
    }
    if (c->any_cx) {
      if (mixed) {
        HIPCHK(c, hipMalloc((void**)&B.ck4f, sizeof(float) * nslot));
        HIPCHK(c, hipMemset(B.ck4f, 0, sizeof(float) * nslot));
      } else {
        HIPCHK(c, hipMalloc((void**)&B.ck4, sizeof(double) * nslot));
        HIPCHK(c, hipMemset(B.ck4, 0, sizeof(double) * nslot));
      }
    }
  }
  c->nkeys_tot = goff + 2LL * nt;
  HIPCHK(c, hipMalloc((void**)&c->d_keycnt, sizeof(unsigned int) * c->nkeys_tot));
  HIPCHK(c, hipMemset(c->d_keycnt, 0, sizeof(unsigned int) * c->nkeys_tot));
  HIPCHK(c, hipMalloc((void**)&c->d_cursor, sizeof(unsigned long long) * c->nkeys_tot));
  HIPCHK(c, hipMalloc((void**)&c->d_cnt_new, sizeof(long long) * nt));
  std::vector<long long> zc(nt, 0);
  HIPCHK(c, up((void**)&c->d_count, zc.data(), sizeof(long long) * nt));
  HIPCHK(c, up((void**)&c->d_next_id, zc.data(), sizeof(long long) * nt));
  HIPCHK(c, hipMalloc((void**)&c->d_hist, sizeof(unsigned int) * (size_t)(hoff ? hoff : 1)));
  HIPCHK(c, hipMalloc((void**)&c->d_nqc, sizeof(unsigned int) * (size_t)goff));
  HIPCHK(c, hipMalloc((void**)&c->d_nqoff, sizeof(unsigned int) * (size_t)goff));
  HIPCHK(c, hipMemset(c->d_hist, 0, sizeof(unsigned int) * (size_t)(hoff ? hoff : 1)));
  HIPCHK(c, hipMemset(c->d_nqc, 0, sizeof(unsigned int) * (size_t)goff));
  HIPCHK(c, hipMemset(c->d_nqoff, 0, sizeof(unsigned int) * (size_t)goff));
  HIPCHK(c, hipMalloc((void**)&c->d_work, sizeof(unsigned int)));
  HIPCHK(c, up((void**)&c->d_live, zc.data(), sizeof(long long) * nt));
  {
    std::vector<int4> bx(nt);
    for (int t = 0; t < nt; ++t) bx[t] = make_int4(1, c->dt[t].nvpar, 1, c->dt[t].nvperp);
    HIPCHK(c, up((void**)&c->d_vbox, bx.data(), sizeof(int4) * nt));
    std::vector<int4> hb;
    for (int t = 0; t < nt; ++t)
      for (int q = 0; q < c->dt[t].nq; ++q) hb.push_back(make_int4(1, c->dt[t].nvpar, 1, c->dt[t].nvperp));
    HIPCHK(c, up((void**)&c->d_hbox, hb.data(), sizeof(int4) * hb.size()));
    HIPCHK(c, hipMalloc((void**)&c->d_vmarg_par, sizeof(unsigned int) * (size_t)voff));
    HIPCHK(c, hipMalloc((void**)&c->d_vmarg_perp, sizeof(unsigned int) * (size_t)woff));
    HIPCHK(c, hipMemset(c->d_vmarg_par, 0, sizeof(unsigned int) * (size_t)voff));
    HIPCHK(c, hipMemset(c->d_vmarg_perp, 0, sizeof(unsigned int) * (size_t)woff));
  }
  HIPCHK(c, hipMalloc((void**)&c->d_stepno, sizeof(long long)));
  HIPCHK(c, hipMemset(c->d_stepno, 0, sizeof(long long)));
  if (c->cfg.flags & KAOS_FLAG_VPERP3D) {
    c->v3d = true;
    c->nhist3_tot = h3off;
    // window of the binning pass: whole config cells, >= 16384 counters where the
    // grid is small, one cell up to 65536 counters (larger grids: global atomics only)
    long long words = 0;
    for (int t = 0; t < nt; ++t) {
      const DevTube& T = c->dt[t];
      const long long nc3 = (long long)T.nv1 * T.nv2 * T.nvpar;
      if (nc3 > 2LL * KAOS_BIN3_MAXWORDS) continue;
      const long long wnq = std::min<long long>(T.nq, std::max<long long>(1, 16384 / nc3));
      words = std::max(words, (wnq * nc3 + 1) / 2);
    }
    c->bin3_words = (int)std::min<long long>(words, KAOS_BIN3_MAXWORDS);
    // binning groups: runs of <= KAOS_BIN3_TILES tiles that never cross a tube,
    // so that every group's LDS window serves all its ions
    std::vector<int2> grp;
    for (int i = 0; i < c->ntiles;) {
      int j = i + 1;
      while (j < c->ntiles && j - i < KAOS_BIN3_TILES && tile_tube[j] == tile_tube[i]) ++j;
      grp.push_back(make_int2(i, j));
      i = j;
    }
    c->nb3 = (int)grp.size();
    HIPCHK(c, up((void**)&c->d_grp3, grp.data(), sizeof(int2) * std::max<size_t>(grp.size(), 1)));
    if (c->bin3_words > 0) {
      HIPCHK(c, hipFuncSetAttribute((const void*)k_bin3d<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(unsigned int) * c->bin3_words)));
      HIPCHK(c, hipFuncSetAttribute((const void*)k_bin3d<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(unsigned int) * c->bin3_words)));
    }
    HIPCHK(c, up((void**)&c->d_v1GH, v1GHx.data(), sizeof(double) * v1off));
    HIPCHK(c, up((void**)&c->d_v1GC, v1GC.data(), sizeof(double) * v1off));
    HIPCHK(c, up((void**)&c->d_dv1, dv1.data(), sizeof(double) * v1off));
    HIPCHK(c, up((void**)&c->d_v2GH, v2GHx.data(), sizeof(double) * v2off));
    HIPCHK(c, up((void**)&c->d_v2GC, v2GC.data(), sizeof(double) * v2off));
    HIPCHK(c, up((void**)&c->d_dv2, dv2.data(), sizeof(double) * v2off));
    HIPCHK(c, up((void**)&c->d_d3v3, d3v3.data(), sizeof(double) * d3off));
    std::vector<double> z8(8 * goff, 0.0);
    HIPCHK(c, up((void**)&c->d_M3, z8.data(), sizeof(double) * 8 * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_hist3, sizeof(unsigned int) * (size_t)h3off));
    HIPCHK(c, hipMemset(c->d_hist3, 0, sizeof(unsigned int) * (size_t)h3off));
  }
  HIPCHK(c, hipMalloc((void**)&c->d_pushed, sizeof(unsigned long long)));
  HIPCHK(c, hipMemset(c->d_pushed, 0, sizeof(unsigned long long)));
  HIPCHK(c, hipMalloc((void**)&c->d_status, sizeof(unsigned int)));
  HIPCHK(c, hipMemset(c->d_status, 0, sizeof(unsigned int)));
  std::vector<unsigned long long> ze(2 * nt, 0ull);
  HIPCHK(c, up((void**)&c->d_esc_acc, ze.data(), sizeof(unsigned long long) * 2 * nt));
  HIPCHK(c, up((void**)&c->d_esc_rep, ze.data(), sizeof(unsigned long long) * 2 * nt));
  if (c->any_cx) {
    std::vector<double> nsn, geo;
    std::vector<long long> eoff(nt + 1, 0);
    for (int t = 0; t < nt; ++t) {
      const HostTube& h = c->ht[t];
      for (int i = 0; i < h.g.nq; ++i) {
        nsn.push_back(h.nsneut.empty() ? 0.0 : h.nsneut[i]);
        geo.push_back(h.cth[i]); geo.push_back(h.ell[i]);
      }
      eoff[t] = c->dt[t].eoff;
    }
    eoff[nt] = c->ena_tot;
    const size_t ne = (size_t)std::max<long long>(c->ena_tot, 1);
    HIPCHK(c, up((void**)&c->d_nsneut, nsn.data(), sizeof(double) * goff));
    HIPCHK(c, up((void**)&c->d_cxgeo, geo.data(), sizeof(double) * 2 * goff));
    HIPCHK(c, up((void**)&c->d_eoff, eoff.data(), sizeof(long long) * (nt + 1)));
    HIPCHK(c, hipMalloc((void**)&c->d_cx, sizeof(double) * 3 * goff));
    HIPCHK(c, hipMemset(c->d_cx, 0, sizeof(double) * 3 * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_cx_need, sizeof(long long) * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_cx_pref, sizeof(long long) * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_cx_thr, sizeof(long long) * goff));
    if ((double)goff * KAOS_CX_BINS >= 4294967296.0)
      return fail(c, KAOS_E_ARG, "charge exchange: too many config cells for the id select's 32-bit keys");
    HIPCHK(c, hipMalloc((void**)&c->d_cx_lev, sizeof(int) * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_cx_hist, sizeof(unsigned int) * KAOS_CX_BINS * (size_t)goff));
    HIPCHK(c, hipMemset(c->d_cx_hist, 0, sizeof(unsigned int) * KAOS_CX_BINS * (size_t)goff));
    HIPCHK(c, hipMalloc((void**)&c->d_ena, sizeof(double) * 6 * ne));
    HIPCHK(c, hipMalloc((void**)&c->d_eid, sizeof(long long) * ne));
    HIPCHK(c, hipMalloc((void**)&c->d_eq, sizeof(int) * ne));
    HIPCHK(c, hipMemset(c->d_eq, 0xFF, sizeof(int) * ne));       // never used: see k_ena_step
    HIPCHK(c, up((void**)&c->d_ena_cnt, zc.data(), sizeof(long long) * nt));
    HIPCHK(c, hipMalloc((void**)&c->d_nq_ena, sizeof(unsigned int) * goff));
    HIPCHK(c, hipMemset(c->d_nq_ena, 0, sizeof(unsigned int) * goff));
    HIPCHK(c, up((void**)&c->d_ena_esc, ze.data(), sizeof(unsigned long long) * 2 * nt));
    HIPCHK(c, up((void**)&c->d_ena_vp, ena_vp.data(), sizeof(double) * ena_vp.size()));
    HIPCHK(c, up((void**)&c->d_ena_vq, ena_vq.data(), sizeof(double) * ena_vq.size()));
    HIPCHK(c, up((void**)&c->d_ena_vf, ena_vf.data(), sizeof(double) * ena_vf.size()));
    HIPCHK(c, up((void**)&c->d_ena_hh, ena_hh.data(), sizeof(double) * ena_hh.size()));
    HIPCHK(c, up((void**)&c->d_ena_d33v, ena_d33.data(), sizeof(double) * ena_d33.size()));
    std::vector<double> z8e(8 * goff, 0.0);
    HIPCHK(c, up((void**)&c->d_ena_M, z8e.data(), sizeof(double) * 8 * goff));
    HIPCHK(c, hipMalloc((void**)&c->d_evc, sizeof(unsigned int) * ne));
    HIPCHK(c, hipMalloc((void**)&c->d_ena_hist, sizeof(unsigned int) * (size_t)std::max(c->ena_hist_tot, 1LL)));
    HIPCHK(c, hipMemset(c->d_ena_hist, 0, sizeof(unsigned int) * (size_t)std::max(c->ena_hist_tot, 1LL)));
  }
  // injection bookkeeping: ions per tube per master step and their prefix
  std::vector<long long> ib(nt + 1, 0), ic(nt, 0);
  for (int t = 0; t < nt; ++t) {
    const DevTube& T = c->dt[t];
    ic[t] = (T.lbcond ? T.n_inj_lb : 0) + (T.ubcond ? T.n_inj_ub : 0);
    ib[t + 1] = ib[t] + ic[t];
  }
  c->inj_total = ib[nt];
  c->inj_base_h = ib;
  c->cnt_h.assign(nt, 0);
  c->live_h.assign(nt, 0);
  HIPCHK(c, up((void**)&c->d_inj_base, ib.data(), sizeof(long long) * (nt + 1)));
  HIPCHK(c, up((void**)&c->d_inj_cnt, ic.data(), sizeof(long long) * nt));
  HIPCHK(c, up((void**)&c->d_uni_base, ib.data(), sizeof(long long) * (nt + 1)));
  c->built = true;
  if (!plain_push(c)) {                         // only the plain push has a 512-thread variant
    std::vector<kaos_ctx::Launch> keep;
    for (const auto& l : c->tune_cand) if (l.block == 256) keep.push_back(l);
    c->tune_cand = keep;
    c->launch.block = 256;
  }
  return KAOS_OK;
}

static void set_state(StepArgs& a, const StateBufs& B) {
  a.rs = B.rs; a.vp = B.vp; a.vr = B.vr; a.qk = B.qk; a.ae = B.ae; a.ag = B.ag; a.vr2 = B.vr2;
  a.rdi = B.rdi; a.vpi = B.vpi; a.lo = B.lo; a.vrf = B.vrf; a.qkf = B.qkf; a.aef = B.aef; a.agf = B.agf; a.vrf2 = B.vrf2;
  a.cell = B.cell; a.id = B.id; a.c12 = B.c12;
  a.ck4 = B.ck4; a.ck4f = B.ck4f;
}

static StepArgs make_args(kaos_ctx* c, int interp, int count, int remove, int qk) {
  StepArgs a;
  a.tubes = c->d_tubes; a.tile_tube = c->d_tile_tube; a.tile_start = c->d_tile_start;
  a.ntiles = c->ntiles;
  a.qGH = c->d_qGH; a.qGC = c->d_qGC; a.EA = c->d_EA; a.EG = c->d_EG;
  a.vparGH = c->d_vparGH; a.vperpGH = c->d_vperpGH;
  set_state(a, c->st[c->cur]);
  a.rGH = c->d_rGH; a.rref = c->d_rref; a.rk = c->d_rk; a.vk = c->d_vk;
  a.lds_nq = c->lds_nq; a.lds_nv = c->lds_nv;
  a.count = c->d_count; a.hist = c->d_hist; a.nqoff = c->d_nqoff;
  a.esc_acc = c->d_esc_acc; a.status = c->d_status; a.pushed = c->d_pushed;
  a.tiles_per_block = c->launch.tpb;
  a.hwords = c->launch.hwords;
  a.persistent = c->launch.persistent;
  a.ngroups = (c->ntiles + c->launch.tpb - 1) / c->launch.tpb;
  a.work = c->d_work;
  a.vbox = c->d_vbox;
  a.stepno = c->d_stepno;
  a.seed = c->cfg.seed;
  a.hbox = c->d_hbox;
  a.do_interp = interp; a.do_count = c->v3d ? 0 : count; a.do_remove = remove; a.do_qk = qk;
  a.v1GH = c->d_v1GH; a.v2GH = c->d_v2GH; a.hist3 = c->d_hist3; a.grp3 = c->d_grp3;
  return a;
}

extern "C" int kaos_upload_vperp(kaos_ctx* c, int32_t tube, int64_t n, const double* vperp1, const double* vperp2) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (c->last_step != 0) return fail(c, KAOS_E_STATE, "kaos_upload_vperp after the first step");
  if (!c->ht[tube].uploaded || n != c->ht[tube].count)
    return fail(c, KAOS_E_STATE, "kaos_upload_vperp needs the tube's uploaded ions (same n)");
  if (n > 0 && (!vperp1 || !vperp2)) return fail(c, KAOS_E_ARG, "bad vperp arrays");
  if (n == 0) return KAOS_OK;
  const DevTube& T = c->dt[tube];
  double *d1 = nullptr, *d2 = nullptr;
  HIPCHK(c, hipMalloc((void**)&d1, sizeof(double) * n));
  HIPCHK(c, hipMalloc((void**)&d2, sizeof(double) * n));
  HIPCHK(c, hipMemcpy(d1, vperp1, sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d2, vperp2, sizeof(double) * n, hipMemcpyHostToDevice));
  StepArgs a = make_args(c, 0, 0, 0, 0);
  hipLaunchKernelGGL(k_set_vperp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, T, (long long)n, d1,
                     d2, c->d_vperpGH + T.woff, a, (int)(c->cfg.precision == KAOS_MIXED));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(d1); (void)hipFree(d2);
  return KAOS_OK;
}

extern "C" int kaos_upload_particles(kaos_ctx* c, int32_t tube, int64_t n, const double* q0,
                                     const double* vpar0, const double* vperp0) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (n < 0 || (n > 0 && (!q0 || !vpar0 || !vperp0))) return fail(c, KAOS_E_ARG, "bad particle arrays");
  if (c->last_step != 0) return fail(c, KAOS_E_STATE, "upload after the first step");
  if (!c->built) { int rc = build_device(c); if (rc) return rc; }
  HostTube& h = c->ht[tube];
  if (n > h.cap) return fail(c, KAOS_E_CAPACITY, "more ions than the tube capacity");
  const DevTube& T = c->dt[tube];
  double *dq0 = nullptr, *dv0 = nullptr, *dr0 = nullptr;
  size_t nb = sizeof(double) * (size_t)(n ? n : 1);
  HIPCHK(c, hipMalloc((void**)&dq0, nb));
  HIPCHK(c, hipMalloc((void**)&dv0, nb));
  HIPCHK(c, hipMalloc((void**)&dr0, nb));
  if (n) {
    HIPCHK(c, hipMemcpy(dq0, q0, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dv0, vpar0, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dr0, vperp0, sizeof(double) * n, hipMemcpyHostToDevice));
    int nbk = (int)((n + 255) / 256);
    StepArgs a = make_args(c, 0, 0, 0, 0);
    hipLaunchKernelGGL(k_upload, dim3(nbk), dim3(256), 0, c->stream, T, (long long)n, dq0, dv0, dr0,
                       c->d_qGH + T.goff, c->d_vparGH + T.voff, c->d_vperpGH + T.woff, c->d_rref + T.goff,
                       a, (int)(c->cfg.precision == KAOS_MIXED));
    HIPCHK(c, hipGetLastError());
  }
  long long cnt = n;
  HIPCHK(c, hipMemcpyAsync(c->d_count + tube, &cnt, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_live + tube, &cnt, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_next_id + tube, &cnt, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(dq0); (void)hipFree(dv0); (void)hipFree(dr0);
  h.count = n;
  h.uploaded = true;
  return KAOS_OK;
}

extern "C" int kaos_cell_counts(int32_t nq, const double* density, const double* d3x, double nsnormfac,
                                int64_t* counts) {
  if (nq < 1 || !density || !d3x || !counts || !(nsnormfac > 0.0)) return KAOS_E_ARG;
  for (int32_t Q = 0; Q < nq; ++Q) {
    int64_t k = kaos_cell_count(density[Q], d3x[Q], nsnormfac);
    if (k < 0) return KAOS_E_ARG;
    counts[Q] = k;
  }
  return KAOS_OK;
}

// ---------------------------------------------------------------------------
// Grid set-up of one tube on the device (kaos_generate_grid).  The shared
// header holds the arithmetic (kaos_gg_*); the kernels below give each cell,
// velocity cell and ENA bin a thread.

struct GGPar {
  kaos_gridgen_in in;
  kaos_gridgen_dims d;
  long long ngp;                 // q grid points NqUB + 3
  double npar, nperp;            // points per side of the symmetric grids (NVGpF / 2)
  double gb_par, gb_perp;        // VparGB, VperpGB (ions)
  double sig_par;                // Vparsigma (the SYMVPARflag = 0 V|| grid starts at -4 sigma)
  double gb_pphi, gb_q, sig_n;   // ENA: VpGB = VphiGB, VqGB, Vqsigma
};

__global__ void k_gg_cells(GGPar P, kaos_gg_cell_t* cells, double* icb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.d.nq + 2) return;
  kaos_gg_cell_t c;
  kaos_gg_cell(P.in.qGA, P.in.qGB, P.ngp, P.in.L, i + 1, &c);
  cells[i] = c;
  icb[i] = kaos_gg_icbbp(&c);
}

// Loop cell Q = set-up cell Q + 1: geometry, temperatures, the hydrostatic
// density (potential sum over set-up cells 1 .. Q + 1 in index order, minus
// the first) and its truncated counts, the neutral counts.
__global__ void k_gg_config(GGPar P, const kaos_gg_cell_t* cells, const double* icb, double* o, long long* counts) {
  const int Q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nq = P.d.nq;
  if (Q >= nq) return;
  const kaos_gg_cell_t c = cells[Q + 1];
  double sum = 0.0;
  for (int k = 0; k <= Q + 1; ++k) sum = sum + icb[k];
  const double gC = sum - icb[0];
  const double ns = kaos_gg_ns(gC, P.in.ms, P.in.ti, P.in.te, P.in.ns0, P.in.qGA);
  const double x = (ns / P.in.nsnormfac) * c.d3x;
  double xn = 0.0;
  if (P.in.cx) xn = (kaos_gg_ns_neut(c.r, P.in.ms, P.in.t_neut, P.in.zns0_neut, P.in.ns0_neut, P.in.qGA) /
                     P.in.nsnormfac) * c.d3x;
  const double v[13] = {c.qGL, c.qGH, c.qGC, c.hq, c.dq, c.d3x, c.r, c.cth, c.ell,
                        (1.0 / 3.0) * P.in.ti + (2.0 / 3.0) * P.in.ti, P.in.te, ns, trunc(xn)};
  for (int k = 0; k < 13; ++k) o[(long long)k * nq + Q] = v[k];
  counts[Q] = (x > -9.2e18 && x < 9.2e18) ? (long long)trunc(x) : -1LL;
}

// Point i (1-based) of velocity axis a: 0 V||, 1 Vperp (2-D), 2 Vperp1,
// 3 Vperp2, 4 ENA Vp, 5 ENA Vq, 6 ENA Vphi.
__device__ double kd_gg_point(const GGPar& P, int a, long long i) {
  const int south = P.in.qGA <= 0.0;
  switch (a) {
    case 0:
      return P.in.symvpar ? kaos_gg_vsym(P.npar, P.gb_par, i)
                          : kaos_gg_vlin((double)P.in.nvpar_gp, -4.0 * P.sig_par, P.gb_par, south, i);
    case 1: case 2: case 3:
      return kaos_gg_vsym(P.nperp, P.gb_perp, i);
    case 5:
      return P.in.symvpar ? kaos_gg_vena(P.npar, P.gb_q, i)
                          : kaos_gg_vlin((double)P.in.nvpar_gp, -4.0 * P.sig_n, P.gb_q, south, i);
    default:
      return kaos_gg_vena(P.nperp, P.gb_pphi, i);
  }
}

// Cells of the seven axes (ncells[a], 0 = axis not generated): edges,
// centre, width into o + 4 * off[a] ([GL | GH | GC | d], n each); V||
// centres of 0 become 1e-15.
__global__ void k_gg_vel(GGPar P, int4 n03, int4 n46, int4 off03, int4 off46, double* o) {
  const int ncell[7] = {n03.x, n03.y, n03.z, n03.w, n46.x, n46.y, n46.z};
  const int off[7] = {off03.x, off03.y, off03.z, off03.w, off46.x, off46.y, off46.z};
  const int a = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = ncell[a];
  if (i >= n) return;
  const double gl = kd_gg_point(P, a, i + 1), gh = kd_gg_point(P, a, i + 2);
  double gc = (gl + gh) / 2.0;
  if (a == 0 && gc == 0.0) gc = 1e-15;
  double* b = o + off[a];
  b[i] = gl; b[n + i] = gh; b[2 * n + i] = gc; b[3 * n + i] = fabs(gh - gl);
}

// Ion d3v: 2-D [Vperp][V||] = 2 pi |Vperp_c| dVperp dV|| (1e-9 if 0); 3-D
// [V1][V2][V||] = dV1 dV2 dV|| (d3vCTp).
__global__ void k_gg_d3v(int three, int n1, int n2, int np, const double* c1, const double* c2, const double* cp,
                         double* d3v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = three ? (long long)n1 * n2 * np : (long long)n1 * np;
  if (i >= n) return;
  const int ip = (int)(i % np);
  const double dpar = cp[3 * np + ip];
  if (three) {
    const int i2 = (int)((i / np) % n2), i1 = (int)(i / ((long long)np * n2));
    d3v[i] = (((1.0 * 1.0) * 1.0 * c1[3 * n1 + i1]) * c2[3 * n2 + i2]) * dpar;
  } else {
    const int i1 = (int)(i / np);
    const double gc = c1[2 * n1 + i1];
    double d = ((((1.0 * 1.0) * fabs(gc)) * c1[3 * n1 + i1]) * dpar) * (2.0 * KAOS_PI);
    if (d == 0.0) d = 1e-9;
    d3v[i] = d;
  }
}

// ENA metric of loop cell Q and bin (ip, iq, iphi), from set-up cell Q's
// theta (cells[Q], 0-based).
__global__ void k_gg_ena(int nq, int nvp, int nvq, int nvphi, const kaos_gg_cell_t* cells, const double* vp,
                         const double* vq, const double* vphi, double* hh, double* d33) {
  const long long nb = (long long)nvp * nvq * nvphi;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * nq) return;
  const int Q = (int)(i / nb);
  const long long b = i % nb;
  const int iphi = (int)(b % nvphi), iq = (int)((b / nvphi) % nvq), ip = (int)(b / ((long long)nvphi * nvq));
  const kaos_gg_cell_t c = cells[Q];
  kaos_gg_ena_metric(c.cth, c.sth, c.ell, vp[2 * nvp + ip], vq[2 * nvq + iq], vp[3 * nvp + ip], vq[3 * nvq + iq],
                     vphi[3 * nvphi + iphi], &hh[i], &d33[i]);
}

extern "C" int kaos_gridgen_dims_of(const kaos_gridgen_in* in, kaos_gridgen_dims* d) {
  if (!in || !d) return KAOS_E_ARG;
  memset(d, 0, sizeof(*d));
  if (in->nq_ub < 11 || in->nq_ub > KAOS_MAXQ) return KAOS_E_ARG;
  if (in->nvpar_gp < 4 || in->nvperp_gp < 4 || (in->nvpar_gp & 1) || (in->nvperp_gp & 1)) return KAOS_E_ARG;
  if (in->nvpar_gp - 2 > 255 || in->nvperp_gp - 2 > 255) return KAOS_E_ARG;
  if (!(in->ms > 0.0) || !(in->ti > 0.0) || !(in->te >= 0.0) || !(in->nsnormfac > 0.0) || !(in->L > 1.0) ||
      !(in->ns0 >= 0.0) || !(in->vpar_sigma_fac > 0.0) || !(in->vperp_sigma_fac > 0.0) ||
      !std::isfinite(in->qGA) || !std::isfinite(in->qGB) || in->qGA == in->qGB)
    return KAOS_E_ARG;
  if (in->cx && (!(in->t_neut > 0.0) || !(in->zns0_neut > 0.0) || !(in->ns0_neut >= 0.0))) return KAOS_E_ARG;
  d->nq = (int32_t)in->nq_ub;
  d->nvpar = (int32_t)(in->nvpar_gp - 2);
  if (in->vperp3d) d->nvperp1 = d->nvperp2 = (int32_t)(in->nvperp_gp - 2);
  else d->nvperp = (int32_t)(in->nvperp_gp - 2);
  if (in->cx) {
    d->nvp = d->nvphi = (int32_t)(in->nvperp_gp - 2);
    d->nvq = (int32_t)(in->nvpar_gp - 2);
  }
  return KAOS_OK;
}

extern "C" int kaos_generate_grid(const kaos_gridgen_in* in, kaos_gridgen_out* out) {
  if (!in || !out) return KAOS_E_ARG;
  GGPar P;
  int rc = kaos_gridgen_dims_of(in, &P.d);
  if (rc) return rc;
  P.in = *in;
  P.ngp = in->nq_ub + 3;
  P.npar = (double)in->nvpar_gp / 2.0;
  P.nperp = (double)in->nvperp_gp / 2.0;
  P.sig_par = kaos_thermal_speed(in->ti, in->ms);
  P.gb_par = in->vpar_sigma_fac * P.sig_par;
  P.gb_perp = in->vperp_sigma_fac * kaos_thermal_speed(in->ti, in->ms);
  P.sig_n = in->cx ? kaos_thermal_speed(in->t_neut, in->ms) : 0.0;        // mNeut = msT
  P.gb_pphi = in->vperp_sigma_fac * P.sig_n;
  P.gb_q = in->vpar_sigma_fac * P.sig_n;
  const kaos_gridgen_dims& d = P.d;
  const int nq = d.nq;
  const int nax[7] = {d.nvpar, d.nvperp, d.nvperp1, d.nvperp2, d.nvp, d.nvq, d.nvphi};
  int off[7], tot = 0;
  for (int a = 0; a < 7; ++a) { off[a] = tot; tot += 4 * nax[a]; }
  const long long n3 = in->vperp3d ? (long long)d.nvperp1 * d.nvperp2 * d.nvpar : (long long)d.nvperp * d.nvpar;
  const long long nena = (long long)nq * d.nvp * d.nvq * d.nvphi;
  // one device block: cells, potential terms, config outputs, velocity cells, d3v, ENA metric, counts
  const size_t ncell0 = (size_t)nq + 2;
  const size_t bytes = ncell0 * sizeof(kaos_gg_cell_t) + ncell0 * 8 + (size_t)13 * nq * 8 + (size_t)tot * 8 +
                       (size_t)n3 * 8 + (size_t)2 * nena * 8 + (size_t)nq * 8;
  if (hipSetDevice(in->device) != hipSuccess) return KAOS_E_DEVICE;
  char* buf = nullptr;
  if (hipMalloc((void**)&buf, bytes) != hipSuccess) return KAOS_E_DEVICE;
  kaos_gg_cell_t* cells = (kaos_gg_cell_t*)buf;
  double* icb = (double*)(buf + ncell0 * sizeof(kaos_gg_cell_t));
  double* cfg = icb + ncell0;
  double* vel = cfg + (size_t)13 * nq;
  double* d3v = vel + tot;
  double* hh = d3v + n3;
  double* d33 = hh + nena;
  long long* cnt = (long long*)(d33 + nena);
  hipLaunchKernelGGL(k_gg_cells, dim3((unsigned)((ncell0 + 255) / 256)), dim3(256), 0, 0, P, cells, icb);
  hipLaunchKernelGGL(k_gg_config, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, 0, P, cells, icb, cfg, cnt);
  hipLaunchKernelGGL(k_gg_vel, dim3(1, 7), dim3(256), 0, 0, P, make_int4(nax[0], nax[1], nax[2], nax[3]),
                     make_int4(nax[4], nax[5], nax[6], 0), make_int4(off[0], off[1], off[2], off[3]),
                     make_int4(off[4], off[5], off[6], 0), vel);
  if (n3 > 0) {
    const int a1 = in->vperp3d ? 2 : 1;
    hipLaunchKernelGGL(k_gg_d3v, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, 0, in->vperp3d ? 1 : 0,
                       nax[a1], in->vperp3d ? nax[3] : 1, nax[0], vel + off[a1], vel + off[3], vel + off[0], d3v);
  }
  if (nena > 0)
    hipLaunchKernelGGL(k_gg_ena, dim3((unsigned)((nena + 255) / 256)), dim3(256), 0, 0, nq, d.nvp, d.nvq, d.nvphi,
                       cells, vel + off[4], vel + off[5], vel + off[6], hh, d33);
  rc = KAOS_OK;
  std::vector<double> hcfg((size_t)13 * nq), hvel(tot > 0 ? tot : 1);
  std::vector<long long> hcnt(nq);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(hcfg.data(), cfg, sizeof(double) * 13 * nq, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hvel.data(), vel, sizeof(double) * tot, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hcnt.data(), cnt, sizeof(long long) * nq, hipMemcpyDeviceToHost) != hipSuccess)
    rc = KAOS_E_DEVICE;
  auto down = [&](double* dst, const double* src, long long n) {
    if (rc == KAOS_OK && dst && n > 0 && hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
      rc = KAOS_E_DEVICE;
  };
  if (in->vperp3d) down(out->d3v3, d3v, n3);
  else down(out->d3v, d3v, n3);
  down(out->ena_hh, hh, nena);
  down(out->ena_d33v, d33, nena);
  (void)hipFree(buf);
  if (rc) return rc;
  double* cdst[13] = {out->qGL, out->qGH, out->qGC, out->hqC, out->dqC, out->d3xC, out->rGC, out->cthGC, out->ellGC,
                      out->Ts, out->Te, out->density, in->cx ? out->ns_neut : nullptr};
  for (int k = 0; k < 13; ++k)
    if (cdst[k]) memcpy(cdst[k], hcfg.data() + (size_t)k * nq, sizeof(double) * nq);
  for (int Q = 0; Q < nq; ++Q)
    if (hcnt[Q] < 0) return KAOS_E_ARG;                   // a count beyond int64 (or not finite)
  if (out->counts) memcpy(out->counts, hcnt.data(), sizeof(long long) * nq);
  out->nsnorm_clb = (double)hcnt[0];
  out->nsnorm_cub = (double)hcnt[nq - 1];
  double* ax[7][4] = {{out->vparGL, out->vparGH, out->vparGC, out->dvpar},
                      {out->vperpGL, out->vperpGH, out->vperpGC, out->dvperp},
                      {out->vperp1GL, out->vperp1GH, out->vperp1GC, out->dvperp1},
                      {out->vperp2GL, out->vperp2GH, out->vperp2GC, out->dvperp2},
                      {out->vpGL, out->vpGH, out->vpGC, out->dvp},
                      {out->vqGL, out->vqGH, out->vqGC, out->dvq},
                      {out->vphiGL, out->vphiGH, out->vphiGC, out->dvphi}};
  for (int a = 0; a < 7; ++a)
    for (int k = 0; k < 4; ++k)
      if (nax[a] > 0 && ax[a][k]) memcpy(ax[a][k], hvel.data() + off[a] + (size_t)k * nax[a], sizeof(double) * nax[a]);
  return KAOS_OK;
}

extern "C" int kaos_seed_particles(kaos_ctx* c, int32_t tube, const int64_t* counts, const double* temperature,
                                   uint64_t seed) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (c->last_step != 0) return fail(c, KAOS_E_STATE, "seed after the first step");
  if (!c->built) { int rc = build_device(c); if (rc) return rc; }
  const DevTube& T = c->dt[tube];
  if (!counts || !temperature) return fail(c, KAOS_E_ARG, "bad counts or temperature array");
  std::vector<long long> coff((size_t)T.nq + 1, 0);
  for (int Q = 0; Q < T.nq; ++Q) {
    if (counts[Q] < 0) return fail(c, KAOS_E_ARG, "negative cell count");
    if (counts[Q] > 0 && !(temperature[Q] > 0.0 && std::isfinite(temperature[Q])))
      return fail(c, KAOS_E_ARG, "temperature of a populated cell must be finite and > 0");
    coff[Q + 1] = coff[Q] + counts[Q];
    if (coff[Q + 1] >= (1LL << 28)) return fail(c, KAOS_E_CAPACITY, "more ions than the tube capacity");
  }
  const long long n = coff[T.nq];
  HostTube& h = c->ht[tube];
  if (n > h.cap) return fail(c, KAOS_E_CAPACITY, "more ions than the tube capacity");
  if (n) {
    long long* dcoff = nullptr;
    double* dtemp = nullptr;
    HIPCHK(c, hipMalloc((void**)&dcoff, sizeof(long long) * coff.size()));
    HIPCHK(c, hipMalloc((void**)&dtemp, sizeof(double) * T.nq));
    HIPCHK(c, hipMemcpy(dcoff, coff.data(), sizeof(long long) * coff.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dtemp, temperature, sizeof(double) * T.nq, hipMemcpyHostToDevice));
    StepArgs a = make_args(c, 0, 0, 0, 0);
    hipLaunchKernelGGL(k_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, T, n, dcoff, dtemp,
                       (unsigned long long)seed, (int)tube, c->d_qGC + T.goff, c->d_qGH + T.goff,
                       c->d_vparGH + T.voff, c->d_vperpGH + T.woff, c->d_rref + T.goff, a,
                       (int)(c->cfg.precision == KAOS_MIXED));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dcoff); (void)hipFree(dtemp);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_count + tube, &n, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_live + tube, &n, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_next_id + tube, &n, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  h.count = n;
  h.uploaded = true;
  return KAOS_OK;
}

static bool is_master(const kaos_ctx* c, long long n) { return n >= 2 && (n % c->cfg.ndatfac) == 0; }

static std::pair<hipEvent_t, hipEvent_t> ev_pair(kaos_ctx* c) {
  hipEvent_t a = nullptr, b = nullptr;
  (void)hipEventCreate(&a); (void)hipEventCreate(&b);
  (void)hipEventRecord(a, c->stream);
  return {a, b};
}

// Dynamic LDS of both push kernels: the histogram window (their grid arrays are static).
static size_t push_lds_bytes(const kaos_ctx* c) { return sizeof(unsigned int) * c->launch.hwords; }

// Moments of every (tube, cell): wavefront-per-cell kernel for tubes with <= 64
// Vpar cells, block-per-cell kernel for the others.
static void launch_moments(kaos_ctx* c, const long long* inj) {
  const int wpb = KAOS_BLOCK / 64;
  if (c->v3d) {
    hipLaunchKernelGGL(k_moments3d, dim3(c->nblk_mom), dim3(KAOS_MOM3_LANES), 0, c->stream, c->d_tubes, c->d_blk_tube,
                       c->d_blk_q, c->d_d3x, c->d_d3v3, c->d_dvpar, c->d_dv1, c->d_dv2, c->d_vparGC, c->d_v1GC,
                       c->d_v2GC, c->d_hist3, c->d_nqoff, c->d_nqc, c->d_NqRT, c->d_M, c->d_M3, c->d_count,
                       c->d_next_id, inj, c->d_live, c->d_esc_acc);
    return;
  }
  if (c->n_cells_wave)
    hipLaunchKernelGGL(k_moments_wave, dim3((c->n_cells_wave + wpb - 1) / wpb), dim3(KAOS_BLOCK), 0, c->stream,
                       c->d_tubes, c->d_blk_tube, c->d_blk_q, c->d_cells_wave, c->n_cells_wave, c->d_d3x, c->d_d3v,
                       c->d_dvpar, c->d_dvperp, c->d_vparGC, c->d_vperpGC, c->d_hist, c->d_nqoff, c->d_nqc,
                       c->d_NqRT, c->d_M, c->d_count, c->d_next_id, inj, c->d_hbox, c->d_live, c->d_esc_acc);
  if (c->n_cells_block)
    hipLaunchKernelGGL(k_moments, dim3(c->n_cells_block), dim3(KAOS_BLOCK), 0, c->stream, c->d_tubes,
                       c->d_blk_tube, c->d_blk_q, c->d_d3x, c->d_d3v, c->d_dvpar, c->d_dvperp, c->d_vparGC,
                       c->d_vperpGC, c->d_hist, c->d_nqoff, c->d_nqc, c->d_NqRT, c->d_M, c->d_count, c->d_next_id,
                       inj, c->d_cells_block, c->d_hbox, c->d_live, c->d_esc_acc);
}

// Velocity boxes of the push's LDS windows from the counts of the master step
// just binned (at the first step and at every reorder; between updates a box
// only steers where counts accumulate, so a stale box costs time, never a
// result).
static void launch_box_update(kaos_ctx* c) {
  if (c->v3d) return;                         // the push bins nothing on the 3-D grid
  const int wpb = KAOS_BLOCK / 64;
  hipLaunchKernelGGL(k_marginals, dim3((c->nblk_mom + wpb - 1) / wpb), dim3(KAOS_BLOCK), 0, c->stream, c->d_tubes,
                     c->d_blk_tube, c->d_blk_q, c->nblk_mom, c->d_hbox, c->d_hist, c->d_vmarg_par, c->d_vmarg_perp);
  hipLaunchKernelGGL(k_vbox, dim3(c->cfg.ntubes), dim3(64), 0, c->stream, c->d_tubes, c->d_vmarg_par,
                     c->d_vmarg_perp, c->d_vbox);
}

// Whether the context's push runs the plain instantiation (256-cell LDS
// tables, no charge exchange, and in the mixed form no heating), the one with
// a 512-thread variant.
static bool plain_push(const kaos_ctx* c) {
  return c->lds_nq <= 256 && !c->any_cx && (c->cfg.precision != KAOS_MIXED || !c->any_icr);
}

static void launch_push(kaos_ctx* c, const StepArgs& a, hipEvent_t tune0 = nullptr, hipEvent_t tune1 = nullptr) {
  int nblocks = a.ngroups;
  const bool wide = c->launch.block == 512 && plain_push(c);
  const int bthreads = wide ? 512 : 256;
  if (a.persistent) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->d_work, 1LL);
    // as many blocks as are resident at the kernel's wavefronts per SIMD (4 SIMDs per CU)
    const int per_cu = (c->cfg.precision == KAOS_MIXED ? KAOS_MIXED_WAVES_EU : KAOS_FP64_WAVES_EU) * 256 / bthreads;
    nblocks = std::min(nblocks, per_cu * c->n_cu);
  }
  std::pair<hipEvent_t, hipEvent_t> ev;
  const bool prof = c->prof_on && !c->in_noalloc;
  if (prof) ev = ev_pair(c);
  if (tune0) (void)hipEventRecord(tune0, c->stream);
  const size_t lds = push_lds_bytes(c);
  if (c->cfg.precision == KAOS_MIXED) {
    // the heating instantiation only when a tube heats (its extra registers
    // never cost the plain push)
    if (wide)
      hipLaunchKernelGGL((k_push_mixed<256, false, 512>), dim3(nblocks), dim3(512), lds, c->stream, a);
    else if (c->lds_nq <= 256 && !c->any_icr && !c->any_cx)
      hipLaunchKernelGGL((k_push_mixed<256, false>), dim3(nblocks), dim3(KAOS_MIXED_BLOCK), lds, c->stream, a);
    else if (c->lds_nq <= 256)
      hipLaunchKernelGGL((k_push_mixed<256, true>), dim3(nblocks), dim3(KAOS_MIXED_BLOCK), lds, c->stream, a);
    else if (!c->any_icr && !c->any_cx)
      hipLaunchKernelGGL((k_push_mixed<KAOS_MAXQ, false>), dim3(nblocks), dim3(KAOS_MIXED_BLOCK), lds, c->stream, a);
    else
      hipLaunchKernelGGL((k_push_mixed<KAOS_MAXQ, true>), dim3(nblocks), dim3(KAOS_MIXED_BLOCK), lds, c->stream, a);
  }
  else if (wide)
    hipLaunchKernelGGL((k_push_fp64<256, false, 512>), dim3(nblocks), dim3(512), lds, c->stream, a);
  else if (c->lds_nq <= 256 && !c->any_cx)
    hipLaunchKernelGGL((k_push_fp64<256, false>), dim3(nblocks), dim3(KAOS_FP64_BLOCK), lds, c->stream, a);
  else if (c->lds_nq <= 256)
    hipLaunchKernelGGL((k_push_fp64<256, true>), dim3(nblocks), dim3(KAOS_FP64_BLOCK), lds, c->stream, a);
  else if (!c->any_cx)
    hipLaunchKernelGGL((k_push_fp64<KAOS_MAXQ, false>), dim3(nblocks), dim3(KAOS_FP64_BLOCK), lds, c->stream, a);
  else
    hipLaunchKernelGGL((k_push_fp64<KAOS_MAXQ, true>), dim3(nblocks), dim3(KAOS_FP64_BLOCK), lds, c->stream, a);
  if (tune1) (void)hipEventRecord(tune1, c->stream);
  if (prof) { (void)hipEventRecord(ev.second, c->stream); c->ev_push.push_back(ev); }
}

// Online autotuner: after two warm-up master steps, the next master steps
// each run the push with one candidate launch shape (tiles per group,
// histogram window, grid or persistent, 256 or 512 threads per block), every
// candidate once in each of KAOS_TUNE_SAMPLES passes; the next master step
// reads the timings and keeps the candidate with the smallest of its times
// (several passes, so that a trend over the first steps does not favour the
// candidates timed last).  Results do not depend on the
// launch shape (integer histograms, independent ions), so tuning steps are
// ordinary steps.
static int tune_before_push(kaos_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (c->tune_skip > 0) {
    --c->tune_skip;
    return KAOS_OK;
  }
  const size_t nc = c->tune_cand.size();
  if (c->tune_i < KAOS_TUNE_SAMPLES * nc) {
    std::pair<hipEvent_t, hipEvent_t> ev;
    HIPCHK(c, hipEventCreate(&ev.first));
    HIPCHK(c, hipEventCreate(&ev.second));
    c->tune_ev.push_back(ev);
    c->launch = c->tune_cand[c->tune_i++ % nc];
    *e0 = ev.first; *e1 = ev.second;
    return KAOS_OK;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<float> best(nc, 0.f);
  for (size_t i = 0; i < c->tune_ev.size(); ++i) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->tune_ev[i].first, c->tune_ev[i].second));
    if (i < nc || ms < best[i % nc]) best[i % nc] = ms;
    (void)hipEventDestroy(c->tune_ev[i].first); (void)hipEventDestroy(c->tune_ev[i].second);
  }
  size_t ib = 0;
  for (size_t i = 1; i < nc; ++i)
    if (best[i] < best[ib]) ib = i;
  c->launch = c->tune_cand[ib];
  c->tune_ev.clear();
  c->tuning = false;
  return KAOS_OK;
}

// Master steps between reorders: cfg.reorder_every, 0 = the default, < 0 = never.
static int reorder_period(const kaos_ctx* c) {
  if (c->cfg.reorder_every != 0) return c->cfg.reorder_every;
  return c->cfg.precision == KAOS_MIXED ? KAOS_REORDER_MIXED : KAOS_REORDER_FP64;
}

// Compact and sort every tube's ions by cell into the other state buffer, swap,
// and bring the host mirror of the slot counts up to date (one synchronisation).
static int enqueue_reorder(kaos_ctx* c) {
  const int nt = c->cfg.ntubes;
  const int mixed = c->cfg.precision == KAOS_MIXED;
  StepArgs a = make_args(c, 0, 0, 0, 0);
  StepArgs b = a;
  set_state(b, c->st[c->cur ^ 1]);
  launch_box_update(c);
  hipLaunchKernelGGL(k_reorder_count, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a, mixed, c->d_keycnt);
  hipLaunchKernelGGL(k_reorder_scan, dim3((nt + 63) / 64), dim3(64), 0, c->stream, c->d_tubes, nt, c->d_keycnt,
                     c->d_cursor, c->d_cnt_new);
  hipLaunchKernelGGL(k_reorder_scatter, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a, b, mixed, c->d_cursor);
  hipLaunchKernelGGL(k_reorder_finish, dim3((nt + 63) / 64), dim3(64), 0, c->stream, nt, c->d_cnt_new, c->d_count);
  HIPCHK(c, hipGetLastError());
  c->cur ^= 1;
  c->masters_since_reorder = 0;
  HIPCHK(c, hipMemcpyAsync(c->cnt_h.data(), c->d_count, sizeof(long long) * nt, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < nt; ++t) c->ht[t].count = c->cnt_h[t];
  return KAOS_OK;
}

// The kernels of step n, in stream order (captured as one graph by kaos_run).
static int enqueue_step_kernels(kaos_ctx* c, long long n, const double* d_uni) {
  const int nt = c->cfg.ntubes;
  const bool master = is_master(c, n);
  const bool first = (n == 1);
  const bool premaster = is_master(c, n + 1);
  hipLaunchKernelGGL(k_step_tick, dim3(1), dim3(1), 0, c->stream, c->d_stepno);
  const int mixed = c->cfg.precision == KAOS_MIXED;
  auto launch_bin3d = [&](const StepArgs& a, int first) {
    const int nb = c->nb3;
    if (mixed)
      hipLaunchKernelGGL(k_bin3d<1>, dim3(nb), dim3(KAOS_BIN3_BLOCK), sizeof(unsigned int) * c->bin3_words,
                         c->stream, a, first, c->bin3_words);
    else
      hipLaunchKernelGGL(k_bin3d<0>, dim3(nb), dim3(KAOS_BIN3_BLOCK), sizeof(unsigned int) * c->bin3_words,
                         c->stream, a, first, c->bin3_words);
  };
  auto zero_hist = [&]() {
    if (c->v3d) {
      hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, c->stream, c->d_hist3, c->nhist3_tot);
      hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, c->stream, c->d_nqoff, c->ncell_tot);
      return;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, c->stream, c->d_hist, c->nhist_tot);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, c->stream, c->d_nqoff, c->ncell_tot);
    hipLaunchKernelGGL(k_hbox_reset, dim3((c->nblk_mom + 255) / 256), dim3(256), 0, c->stream, c->nblk_mom,
                       c->d_tubes, c->d_blk_tube, c->d_blk_q, c->d_vbox, c->d_hbox);
  };
  // charge exchange: counts of a master step, ENA outputs reset, conversions, ENA push
  const int ncell = (int)c->ncell_tot;
  auto cx_freq = [&]() {
    hipLaunchKernelGGL(k_cx_freq, dim3(nt), dim3(256), 0, c->stream, c->d_tubes, c->d_M3, c->d_NqRT, c->d_nqc,
                       c->d_d3x, c->d_cxgeo, c->d_nsneut, c->d_cx, c->d_stepno, (unsigned long long)c->cfg.seed);
  };
  auto ena_reset = [&]() {
    const int nr = std::max(ncell, 2 * nt);
    hipLaunchKernelGGL(k_ena_reset, dim3((nr + 255) / 256), dim3(256), 0, c->stream, ncell, nt, c->d_nq_ena,
                       c->d_ena_esc);
    hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, c->stream, c->d_ena_hist, c->ena_hist_tot);
  };
  auto ena_moments = [&]() {
    hipLaunchKernelGGL(k_ena_moments, dim3(ncell), dim3(KAOS_ENA_LANES), 0, c->stream, c->d_tubes, c->d_blk_tube,
                       c->d_blk_q, c->d_nq_ena, c->d_ena_hist, c->d_ena_vp, c->d_ena_vq, c->d_ena_vf, c->d_ena_hh,
                       c->d_ena_d33v, c->d_d3x, c->d_ena_M);
  };
  auto cx_convert = [&]() {
    StepArgs a = make_args(c, 0, 0, 0, 0);
    hipLaunchKernelGGL(k_cx_sel_init, dim3((ncell + 255) / 256), dim3(256), 0, c->stream, ncell, c->d_tubes,
                       c->d_blk_tube, c->d_blk_q, c->d_cx, c->d_next_id, c->d_cx_need, c->d_cx_pref, c->d_cx_thr,
                       c->d_cx_lev);
    for (int L = 0; L < KAOS_CX_LEVELS; ++L) {
      hipLaunchKernelGGL(k_cx_hist, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a, mixed, L, c->d_cx_need,
                         c->d_cx_pref, c->d_cx_lev, c->d_cx_hist);
      hipLaunchKernelGGL(k_cx_scan, dim3((ncell + 3) / 4), dim3(256), 0, c->stream, ncell, L, c->d_cx_lev,
                         c->d_cx_need, c->d_cx_pref, c->d_cx_thr, c->d_cx_hist);
    }
    hipLaunchKernelGGL(k_cx_convert, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a, mixed, c->d_cx_thr,
                       c->d_ena, c->d_eid, c->d_eq, c->d_evc, c->d_ena_cnt, c->ena_tot, c->d_live);
  };
  auto ena_step = [&](int m) {
    if (c->ena_tot > 0)
      hipLaunchKernelGGL(k_ena_step, dim3((unsigned)((c->ena_tot + 255) / 256)), dim3(256), 0, c->stream,
                         c->d_tubes, nt, c->d_eoff, c->d_ena_cnt, c->ena_tot, c->d_ena, c->d_eq, c->d_qGH,
                         c->d_nq_ena, c->d_ena_esc, m, c->d_evc, c->d_ena_vp, c->d_ena_vq, c->d_ena_vf,
                         c->d_ena_hist);
  };
  if (first) {
    StepArgs a = make_args(c, 1, 0, 0, premaster);
    zero_hist();
    if (c->v3d)
      launch_bin3d(a, 1);
    else
      hipLaunchKernelGGL(k_bin_initial, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a);
    // field of master step 1 from "master step 0" (no moments: M = 0), then moments of step 1
    hipLaunchKernelGGL(k_field, dim3(nt), dim3(256), 0, c->stream, c->d_tubes, c->d_hq, c->d_dq, c->d_Te,
                       c->d_M, c->d_M0F, c->d_M1F, c->d_EA, c->d_EAin, c->d_esc_acc, c->d_esc_rep, 0);
    launch_moments(c, nullptr);
    if (c->any_cx) { ena_reset(); ena_moments(); cx_freq(); }
    launch_box_update(c);
    if (c->any_icr && c->cfg.precision != KAOS_MIXED)
      hipLaunchKernelGGL(k_icr, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a);
    launch_push(c, a);
  } else {
    if (c->any_cx) {
      cx_convert();
      if (master) ena_reset();
      ena_step(master ? 1 : 0);
      if (master) ena_moments();
    }
    if (master) {
      hipLaunchKernelGGL(k_field, dim3(nt), dim3(256), 0, c->stream, c->d_tubes, c->d_hq, c->d_dq, c->d_Te,
                         c->d_M, c->d_M0F, c->d_M1F, c->d_EA, c->d_EAin, c->d_esc_acc, c->d_esc_rep, 1);
      zero_hist();
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->tuning && master && !c->in_noalloc) {
      int rc = tune_before_push(c, &e0, &e1);
      if (rc) return rc;
    }
    StepArgs a = make_args(c, master, master, master, premaster);
    if (master && c->v3d)
      launch_bin3d(a, 0);
    if (c->any_icr && c->cfg.precision != KAOS_MIXED)
      hipLaunchKernelGGL(k_icr, dim3(c->ntiles), dim3(KAOS_BLOCK), 0, c->stream, a);
    launch_push(c, a, e0, e1);
    if (master) {
      if (c->inj_total > 0) {
        int nb = (int)((c->inj_total + 255) / 256);
        hipLaunchKernelGGL(k_inject, dim3(nb), dim3(256), 0, c->stream, a, c->d_inj_base, c->d_next_id, nt,
                           c->d_vperpGH, d_uni, c->d_uni_base, (unsigned long long)c->cfg.seed,
                           c->d_stepno, c->inj_total, mixed);
      }
      launch_moments(c, c->d_inj_cnt);
      if (c->any_cx) cx_freq();
    }
  }
  HIPCHK(c, hipGetLastError());
  return KAOS_OK;
}

// Host bookkeeping after the kernels of step n: slot counts, the reorder.
static int finish_step(kaos_ctx* c, long long n) {
  const int nt = c->cfg.ntubes;
  const bool master = is_master(c, n);
  const bool first = (n == 1);
  // host mirror of the slot counts
  if (master)
    for (int t = 0; t < nt; ++t) c->ht[t].count += c->inj_base_h[t + 1] - c->inj_base_h[t];
  if (master && !first && reorder_period(c) > 0 && ++c->masters_since_reorder >= reorder_period(c)) {
    int rc = enqueue_reorder(c);
    if (rc) return rc;
  }
  c->last_step = n;
  c->last_master = first ? 1 : (int)master;
  return KAOS_OK;
}

// Enqueue step n eagerly (no host synchronisation but at a reorder).
static int enqueue_step(kaos_ctx* c, long long n, const double* d_uni) {
  int rc = enqueue_step_kernels(c, n, d_uni);
  if (rc) return rc;
  return finish_step(c, n);
}

// Enqueue step n as a replay of a captured graph of its kernels.  The graph
// depends on the step kind (master, next-step master), the live state buffer
// and the launch shape; it is captured the first time that combination occurs.
static int enqueue_step_graph(kaos_ctx* c, long long n) {
  unsigned long long key = (unsigned long long)is_master(c, n) | ((unsigned long long)is_master(c, n + 1) << 1) |
                           ((unsigned long long)c->cur << 2) | ((unsigned long long)c->launch.persistent << 3) |
                           ((unsigned long long)c->launch.tpb << 4) | ((unsigned long long)c->launch.hwords << 20) |
                           ((unsigned long long)(c->launch.block == 512) << 40);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step_kernels(c, n, nullptr);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    HIPCHK(c, e);
    HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    it = c->graphs.emplace(key, ge).first;
  }
  HIPCHK(c, hipGraphLaunch(it->second, c->stream));
  return finish_step(c, n);
}

static int check_next(kaos_ctx* c, long long n) {
  if (!c->built) return fail(c, KAOS_E_STATE, "no particles uploaded");
  for (int t = 0; t < c->cfg.ntubes; ++t)
    if (!c->ht[t].uploaded) return fail(c, KAOS_E_STATE, "tube " + std::to_string(t) + " not uploaded");
  if (n != c->last_step + 1) return fail(c, KAOS_E_STATE, "steps must be consecutive starting at 1");
  if (is_master(c, n))
    for (int t = 0; t < c->cfg.ntubes; ++t) {
      long long add = c->inj_base_h[t + 1] - c->inj_base_h[t];
      if (c->ht[t].count + add > c->ht[t].cap)
        return fail(c, KAOS_E_CAPACITY, "tube " + std::to_string(t) + " capacity exhausted at step " + std::to_string(n));
    }
  return KAOS_OK;
}

// Error bits the kernels set: a NaN in the state, or a capacity the device
// ran out of (ENA slots, 40-bit ids of the charge-exchange selection).
static int read_status(kaos_ctx* c, int32_t* st) {
  unsigned int s = 0;
  HIPCHK(c, hipMemcpy(&s, c->d_status, sizeof(s), hipMemcpyDeviceToHost));
  *st = (s & (1u << KAOS_E_NAN)) ? KAOS_E_NAN : (s & (1u << KAOS_E_CAPACITY)) ? KAOS_E_CAPACITY : KAOS_OK;
  return KAOS_OK;
}

static const char* status_msg(int32_t st) {
  return st == KAOS_E_NAN ? "NaN in the particle state"
                          : "device capacity exceeded (ENA slots of a tube, or ion ids beyond 2^40)";
}

extern "C" int kaos_step(kaos_ctx* c, int64_t n, const double* uniforms, int64_t n_uniforms,
                         kaos_step_out* out) {
  if (!c) return KAOS_E_ARG;
  int rc = check_next(c, n);
  if (rc) return rc;
  const double* d_uni = nullptr;
  if (uniforms && is_master(c, n) && c->inj_total > 0) {
    if (n_uniforms < 6 * c->inj_total) return fail(c, KAOS_E_ARG, "too few uniforms for this step's injection");
    if (c->uni_cap < 6 * c->inj_total) {
      if (c->d_uni) (void)hipFree(c->d_uni);
      HIPCHK(c, hipMalloc((void**)&c->d_uni, sizeof(double) * 6 * c->inj_total));
      c->uni_cap = 6 * c->inj_total;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_uni, uniforms, sizeof(double) * 6 * c->inj_total, hipMemcpyHostToDevice, c->stream));
    d_uni = c->d_uni;
  }
  // the step's kernels as a replay of its captured graph (as kaos_run does),
  // except with caller uniforms, profiling or tuning on, and at step 1
  const bool graph = !d_uni && !(c->cfg.flags & KAOS_FLAG_NO_GRAPH) && !c->prof_on && !c->tuning && n > 1;
  rc = graph ? enqueue_step_graph(c, n) : enqueue_step(c, n, d_uni);
  if (rc) return rc;
  if (out || d_uni) HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) {
    memset(out, 0, sizeof(*out));
    out->master = c->last_master;
    out->n_injected = is_master(c, n) ? c->inj_total : 0;
    HIPCHK(c, hipMemcpyAsync(c->live_h.data(), c->d_live, sizeof(long long) * c->live_h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t alive = 0;
    for (long long k : c->live_h) alive += k;
    out->n_alive = alive;
    rc = read_status(c, &out->status);
    if (rc) return rc;
    if (out->status) return fail(c, out->status, status_msg(out->status));
  }
  return KAOS_OK;
}

// kaos_step without allocations: the caller's uniforms go to the device's
// `scratch` array, the step is enqueued eagerly (no graph to instantiate), the
// autotuner and kernel timing do not run in it (the launch shape stays), and
// host mirrors sized when the context was built stand in for temporaries.
extern "C" int kaos_step_noalloc(kaos_ctx* c, int64_t n, const double* uniforms, int64_t n_uniforms,
                                 kaos_step_out* out) {
  if (!c) return KAOS_E_ARG;
  int rc = check_next(c, n);
  if (rc) return rc;
  const double* d_uni = nullptr;
  if (uniforms && is_master(c, n) && c->inj_total > 0) {
    const long long need = 6 * c->inj_total;
    if (need > KAOS_NOALLOC_UNIFORMS)
      return fail(c, KAOS_E_ARG, "kaos_step_noalloc: more injected ions per step than its scratch holds "
                                 "(65536); use kaos_step");
    if (n_uniforms < need) return fail(c, KAOS_E_ARG, "too few uniforms for this step's injection");
    double* sp = g_scratch[c->cfg.device];
    if (!sp) return fail(c, KAOS_E_STATE, "no-allocation scratch not set up on this device");
    HIPCHK(c, hipMemcpyAsync(sp, uniforms, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    d_uni = sp;
  }
  c->in_noalloc = true;
  rc = enqueue_step(c, n, d_uni);
  c->in_noalloc = false;
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) {
    memset(out, 0, sizeof(*out));
    out->master = c->last_master;
    out->n_injected = is_master(c, n) ? c->inj_total : 0;
    HIPCHK(c, hipMemcpyAsync(c->live_h.data(), c->d_live, sizeof(long long) * c->live_h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t alive = 0;
    for (long long k : c->live_h) alive += k;
    out->n_alive = alive;
    rc = read_status(c, &out->status);
    if (rc) return rc;
    if (out->status) return fail(c, out->status, status_msg(out->status));
  }
  return KAOS_OK;
}

extern "C" int kaos_run(kaos_ctx* c, int64_t n_first, int64_t nsteps, int64_t* particle_steps) {
  if (!c) return KAOS_E_ARG;
  unsigned long long p0 = 0, p1 = 0;
  HIPCHK(c, hipMemcpy(&p0, c->d_pushed, sizeof(p0), hipMemcpyDeviceToHost));
  std::pair<hipEvent_t, hipEvent_t> ev;
  if (c->prof_on) ev = ev_pair(c);
  for (int64_t i = 0; i < nsteps; ++i) {
    const long long n = n_first + i;
    int rc = check_next(c, n);
    if (rc) return rc;
    const bool graph = !(c->cfg.flags & KAOS_FLAG_NO_GRAPH) && !c->prof_on && !c->tuning && n > 1;
    rc = graph ? enqueue_step_graph(c, n) : enqueue_step(c, n, nullptr);
    if (rc) return rc;
  }
  if (c->prof_on) { (void)hipEventRecord(ev.second, c->stream); c->ev_step.push_back(ev); }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&p1, c->d_pushed, sizeof(p1), hipMemcpyDeviceToHost));
  if (particle_steps) *particle_steps = (int64_t)(p1 - p0);
  int32_t st = 0;
  int rc = read_status(c, &st);
  if (rc) return rc;
  if (st) return fail(c, st, status_msg(st));
  return KAOS_OK;
}

extern "C" int kaos_count(kaos_ctx* c, int32_t tube, int64_t* n) {
  if (!c || !n) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built) { *n = 0; return KAOS_OK; }
  long long k = 0;
  HIPCHK(c, hipMemcpyAsync(&k, c->d_live + tube, sizeof(k), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *n = (int64_t)k;
  return KAOS_OK;
}

extern "C" int kaos_download_master(kaos_ctx* c, int32_t tube, int64_t* nq_counts, int64_t* nph_counts,
                                    double* NqRT, double* M, double* EA, double* EG, double* esc) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built) return fail(c, KAOS_E_STATE, "nothing computed yet");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const HostTube& h = c->ht[tube];
  int nq = T.nq;
  long long ncell = (long long)T.nvpar * T.nvperp;
  if (nq_counts) {
    std::vector<unsigned int> u(nq);
    HIPCHK(c, hipMemcpy(u.data(), c->d_nqc + T.goff, sizeof(unsigned int) * nq, hipMemcpyDeviceToHost));
    for (int i = 0; i < nq; ++i) nq_counts[i] = u[i];
  }
  if (nph_counts) {
    std::vector<unsigned int> u((size_t)(nq * ncell));
    HIPCHK(c, hipMemcpy(u.data(), c->d_hist + T.hoff, sizeof(unsigned int) * nq * ncell, hipMemcpyDeviceToHost));
    for (long long i = 0; i < nq * ncell; ++i) nph_counts[i] = u[i];
  }
  if (NqRT) HIPCHK(c, hipMemcpy(NqRT, c->d_NqRT + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  if (M) HIPCHK(c, hipMemcpy(M, c->d_M + 6 * T.goff, sizeof(double) * 6 * nq, hipMemcpyDeviceToHost));
  if (EA) HIPCHK(c, hipMemcpy(EA, c->d_EA + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  if (EG) HIPCHK(c, hipMemcpy(EG, c->d_EG + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  if (esc) {
    unsigned long long e[2];
    HIPCHK(c, hipMemcpy(e, c->d_esc_acc + 2 * tube, sizeof(e), hipMemcpyDeviceToHost));
    const kaos_grid& g = h.g;
    int ub = nq - 1;
    double dtm = (double)c->cfg.ndatfac * g.h;
    double NLB = (double)e[0] * g.nsnormfac, NUB = (double)e[1] * g.nsnormfac;
    esc[0] = (double)e[0]; esc[1] = (double)e[1];
    esc[2] = NLB / ((h.d3x[0] / (h.hq[0] * h.dq[0])) * dtm);
    esc[3] = NUB / ((h.d3x[ub] / (h.hq[ub] * h.dq[ub])) * dtm);
  }
  return KAOS_OK;
}

extern "C" int kaos_download_master3d(kaos_ctx* c, int32_t tube, int64_t* n3_counts, double* M) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built) return fail(c, KAOS_E_STATE, "nothing computed yet");
  if (!c->v3d) return fail(c, KAOS_E_STATE, "the context has no 3-D velocity grid (KAOS_FLAG_VPERP3D)");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  if (n3_counts) {
    const size_t n = (size_t)T.nq * T.nv1 * T.nv2 * T.nvpar;
    std::vector<unsigned int> u(n);
    HIPCHK(c, hipMemcpy(u.data(), c->d_hist3 + T.h3off, sizeof(unsigned int) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) n3_counts[i] = u[i];
  }
  if (M) HIPCHK(c, hipMemcpy(M, c->d_M3 + 8 * T.goff, sizeof(double) * 8 * T.nq, hipMemcpyDeviceToHost));
  return KAOS_OK;
}

// One stream-unformatted file of DataExport2Sub: open(status='replace',
// form='unformatted', access='stream') + write + close writes the values'
// bytes and nothing else.
static bool export_raw(const std::string& path, const void* p, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  size_t w = bytes ? fwrite(p, 1, bytes, f) : 0;
  bool ok = (w == bytes);
  if (fclose(f) != 0) ok = false;
  return ok;
}

extern "C" int kaos_export_master(kaos_ctx* c, int32_t tube, const kaos_export_info* info, int64_t* nn_out) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!info || !info->dir || !info->theta || !info->Ts || info->species < 1 || info->flux_tube < 1)
    return fail(c, KAOS_E_ARG, "export info needs dir, theta, Ts and 1-based species/flux_tube");
  const long long n = c->last_step;
  if (!c->built || !(n == 1 || is_master(c, n)))
    return fail(c, KAOS_E_STATE, "export after step 1 or a master step only");
  const long long nn = (n == 1) ? 1 : n / c->cfg.ndatfac + 1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const HostTube& h = c->ht[tube];
  const int nq = T.nq, nvpar = T.nvpar, nvperp = T.nvperp;
  const long long ncell = (long long)nvpar * nvperp;
  std::vector<double> M(6 * (size_t)nq), M0F(nq), M1F(nq), EA(nq), EG(nq);
  const bool v3d = c->v3d;
  const long long ncell3 = v3d ? (long long)T.nv1 * T.nv2 * nvpar : 0;
  std::vector<unsigned int> hist(v3d ? 0 : (size_t)(nq * ncell)), hist3((size_t)(nq * ncell3));
  std::vector<double> M3(v3d ? 8 * (size_t)nq : 0);
  HIPCHK(c, hipMemcpy(M.data(), c->d_M + 6 * T.goff, sizeof(double) * 6 * nq, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(M0F.data(), c->d_M0F + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(M1F.data(), c->d_M1F + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(EA.data(), c->d_EA + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(EG.data(), c->d_EG + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  if (v3d) {
    HIPCHK(c, hipMemcpy(hist3.data(), c->d_hist3 + T.h3off, sizeof(unsigned int) * hist3.size(),
                        hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(M3.data(), c->d_M3 + 8 * T.goff, sizeof(double) * 8 * nq, hipMemcpyDeviceToHost));
  } else {
    HIPCHK(c, hipMemcpy(hist.data(), c->d_hist + T.hoff, sizeof(unsigned int) * hist.size(), hipMemcpyDeviceToHost));
  }

  std::string dir(info->dir);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const std::string pre = dir + std::to_string(nn) + "_" + std::to_string(info->species) + "_" +
                          std::to_string(info->flux_tube) + "_";
  bool ok = true;
  auto put = [&](const std::string& name, const void* p, size_t bytes) {
    if (ok && !export_raw(pre + name + "fort.bin", p, bytes)) ok = false;
  };
  auto put_cells = [&](const char* name, const double* v) { put(name, v, sizeof(double) * nq); };
  // _8_0_2_DataExport2.f90:29-107: time, cadence, reference density, grid rows
  const double time = (n == 1) ? 0.0 : (double)n * h.g.h;       // Time(1) of step n (_8_KineticSolver.f90:178-184)
  const int64_t ndat = c->cfg.ndatfac;
  put("TimeT", &time, sizeof(double));
  put("ndatfacGT", &ndat, sizeof(int64_t));
  put("ns0GT", &info->ns0, sizeof(double));
  const std::vector<double> pGC(nq, h.g.p), phiGC(nq, info->phi), zeros(nq, 0.0);
  put_cells("qGCGT", h.qGC.data());
  put_cells("pGCGT", pGC.data());
  put_cells("rGCGT", h.rGC.data());
  put_cells("phiGCGT", phiGC.data());
  put_cells("thetaGCGT", info->theta);
  put_cells("d3xCGT", h.d3x.data());
  put_cells("TsGT", info->Ts);
  // :167-191: NphRTp(:, :, nn) of every cell, Vperp fastest; NphRT =
  // NphReNormRT (nsnormfac / d3x), 0 for an empty bin (_8_3_1_IonParticleCounts.f90:713-729)
  std::vector<double> nph(v3d ? 0 : ncell);
  for (int Q = 0; Q < nq && ok && !v3d; ++Q) {
    const double w = h.g.nsnormfac / h.d3x[Q];
    for (int i = 0; i < nvperp; ++i)
      for (int b = 0; b < nvpar; ++b) {
        const unsigned int k = hist[(size_t)Q * ncell + (size_t)i * nvpar + b];
        nph[(size_t)b * nvperp + i] = (k == 0u) ? 0.0 : (double)k * w;
      }
    if (!export_raw(pre + std::to_string(Q + 1) + "_NphRTfort.bin", nph.data(), sizeof(double) * ncell)) ok = false;
  }
  // :124-165 (3-D grid): per cell TeNT(nn) (constant in time: dNTe = 0,
  // _8_KineticSolver.f90:66-84), N2PerpphRTp(:, :, :, nn) and d3vCTp(:, :, :),
  // both (Vperp1, Vperp2, Vpar) with Vperp1 fastest; N2PerpphRT = count
  // (nsnormfac / d3x), 0 for an empty bin (_8_3_1_IonParticleCounts.f90:440-457)
  if (v3d) {
    const int n1 = T.nv1, n2 = T.nv2;
    std::vector<double> n3((size_t)ncell3), d3((size_t)ncell3);
    for (int i1 = 0; i1 < n1; ++i1)
      for (int i2 = 0; i2 < n2; ++i2)
        for (int b = 0; b < nvpar; ++b)
          d3[((size_t)b * n2 + i2) * n1 + i1] = h.d3v3[((size_t)i1 * n2 + i2) * nvpar + b];
    for (int Q = 0; Q < nq && ok; ++Q) {
      const double w = h.g.nsnormfac / h.d3x[Q];
      for (int i1 = 0; i1 < n1; ++i1)
        for (int i2 = 0; i2 < n2; ++i2)
          for (int b = 0; b < nvpar; ++b) {
            const unsigned int k = hist3[(size_t)Q * ncell3 + ((size_t)i1 * n2 + i2) * nvpar + b];
            n3[((size_t)b * n2 + i2) * n1 + i1] = (k == 0u) ? 0.0 : (double)k * w;
          }
      const std::string pq = pre + std::to_string(Q + 1) + "_";
      if (!export_raw(pq + "TeNTfort.bin", &h.Te[Q], sizeof(double)) ||
          !export_raw(pq + "N2PerpphRTfort.bin", n3.data(), sizeof(double) * ncell3) ||
          !export_raw(pq + "d3vCTpfort.bin", d3.data(), sizeof(double) * ncell3))
        ok = false;
    }
  }
  // :225-510: moments (3-D grid: M1Perp1/M1Perp2 and M2Perp1/M2Perp2 in place
  // of M1Perp and M2Perp, :279-305, 468-500); :359-410: filtered moments
  // (only M0 and M1Par are set without charge exchange)
  put_cells("M0phRT", M.data());
  if (v3d) {
    put_cells("M1Perp1phRT", M3.data() + nq);
    put_cells("M1Perp2phRT", M3.data() + 2 * nq);
  } else {
    put_cells("M1PerpphRT", M.data() + nq);
  }
  put_cells("M1ParphRT", M.data() + 2 * nq);
  put_cells("M0FiltAvrgRT", M0F.data());
  put_cells("M1Perp1FiltAvrgRT", zeros.data());
  put_cells("M1Perp2FiltAvrgRT", zeros.data());
  put_cells("M1ParFiltAvrgRT", M1F.data());
  put_cells("M2ParFiltAvrgRT", zeros.data());
  put_cells("M2phRT", M.data() + 3 * nq);
  if (v3d) {
    put_cells("M2Perp1phRT", M3.data() + 5 * nq);
    put_cells("M2Perp2phRT", M3.data() + 6 * nq);
  } else {
    put_cells("M2PerpphRT", M.data() + 4 * nq);
  }
  put_cells("M2ParphRT", M.data() + 5 * nq);
  // :645-704: fields.  EAINERTIALflag off: EAInertialRT = abs(0d0), and
  // EAmagRT = EAInertialRT + EAPressureRT (_8_7_AmbipolarEfield.f90:153-155,
  // 1296-1299); EPARflag off: PhiParRT = 0 and EPmagRT = 0 (_8_8_PotentialStructure.f90:46-96)
  std::vector<double> EAmag(nq);
  for (int Q = 0; Q < nq; ++Q) EAmag[Q] = 0.0 + EA[Q];
  put_cells("EAInertialRT", zeros.data());
  put_cells("EAPressureRT", EA.data());
  put_cells("EAmagRT", EAmag.data());
  put_cells("EGmagRT", EG.data());
  put_cells("EPmagRT", zeros.data());
  put_cells("PhiParRT", zeros.data());
  if (!ok) return fail(c, KAOS_E_IO, "cannot write the export files under " + dir);
  if (nn_out) *nn_out = nn;
  return KAOS_OK;
}

extern "C" int kaos_export_spinup(kaos_ctx* c, int32_t tube, const kaos_spinup_info* info) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!info || !info->dir || info->species < 1 || info->flux_tube < 1)
    return fail(c, KAOS_E_ARG, "spin-up info needs dir and 1-based species/flux_tube");
  const long long n = c->last_step;
  if (!c->built || !(n == 1 || is_master(c, n)))
    return fail(c, KAOS_E_STATE, "spin-up export after step 1 or a master step only");
  if (!c->v3d)
    return fail(c, KAOS_E_STATE, "the spin-up temperature needs the 3-D moments (KAOS_FLAG_VPERP3D)");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const int nq = T.nq;
  std::vector<double> M3(8 * (size_t)nq), EA(nq);
  HIPCHK(c, hipMemcpy(M3.data(), c->d_M3 + 8 * T.goff, sizeof(double) * 8 * nq, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(EA.data(), c->d_EA + T.goff, sizeof(double) * nq, hipMemcpyDeviceToHost));
  // TemperatureOutputRT, evaluated left to right as the reference's statement:
  // ((2/3) M2Par / kB + (2/3) M2Perp1 / kB) + (2/3) M2Perp2 / kB
  std::vector<double> Tout(nq), EAmag(nq), zeros(nq, 0.0);
  const double c23 = 2.0 / 3.0;
  for (int Q = 0; Q < nq; ++Q) {
    const double m2par = M3[7 * (size_t)nq + Q], m2p1 = M3[5 * (size_t)nq + Q], m2p2 = M3[6 * (size_t)nq + Q];
    Tout[Q] = ((c23 * m2par) / KAOS_KB + (c23 * m2p1) / KAOS_KB) + (c23 * m2p2) / KAOS_KB;
    EAmag[Q] = 0.0 + EA[Q];
  }
  std::string dir(info->dir);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const std::string pre = dir + std::to_string(info->species) + "_" + std::to_string(info->flux_tube) + "_";
  bool ok = true;
  auto put = [&](const char* name, const void* p, size_t bytes) {
    if (ok && !export_raw(pre + name + "fort.bin", p, bytes)) ok = false;
  };
  put("NqLBT", &info->nq_lb, sizeof(int64_t));
  put("NqUBT", &info->nq_ub, sizeof(int64_t));
  put("DensityOutputRT", M3.data(), sizeof(double) * nq);
  put("TemperatureOutputRT", Tout.data(), sizeof(double) * nq);
  put("EAInertialOutputRT", zeros.data(), sizeof(double) * nq);
  put("EAPressureOutputRT", EA.data(), sizeof(double) * nq);
  put("EAmagOutputRT", EAmag.data(), sizeof(double) * nq);
  put("nsnormCLBOutput", &info->nsnorm_clb, sizeof(double));
  put("nsnormCUBOutput", &info->nsnorm_cub, sizeof(double));
  put("LBREPLENISHflagTOutput", &info->lb_replenish, sizeof(int64_t));
  put("UBREPLENISHflagTOutput", &info->ub_replenish, sizeof(int64_t));
  if (!ok) return fail(c, KAOS_E_IO, "cannot write the spin-up files under " + dir);
  return KAOS_OK;
}

// Whole-file read of a spin-up file that must hold exactly `bytes` bytes.
static bool read_raw(const std::string& path, void* p, size_t bytes) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  size_t r = bytes ? fread(p, 1, bytes, f) : 0;
  bool ok = (r == bytes);
  char extra;
  if (ok && fread(&extra, 1, 1, f) != 0) ok = false;   // longer than its name implies
  fclose(f);
  return ok;
}

extern "C" int kaos_read_spinup(const char* dir, int32_t species, int32_t flux_tube, int64_t nq_cap,
                                kaos_spinup_state* st, double* density, double* temperature, double* ea_inertial,
                                double* ea_pressure, double* ea_mag) {
  if (!dir || !st || species < 1 || flux_tube < 1) {
    g_init_err = "kaos_read_spinup needs dir, st and 1-based species/flux_tube";
    return KAOS_E_ARG;
  }
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  const std::string pre = d + std::to_string(species) + "_" + std::to_string(flux_tube) + "_";
  std::string bad;
  auto get = [&](const char* name, void* p, size_t bytes) {
    if (bad.empty() && !read_raw(pre + name + "fort.bin", p, bytes)) bad = pre + name + "fort.bin";
  };
  kaos_spinup_state S;
  memset(&S, 0, sizeof(S));
  get("NqLBT", &S.nq_lb, sizeof(int64_t));
  get("NqUBT", &S.nq_ub, sizeof(int64_t));
  get("nsnormCLBOutput", &S.nsnorm_clb, sizeof(double));
  get("nsnormCUBOutput", &S.nsnorm_cub, sizeof(double));
  get("LBREPLENISHflagTOutput", &S.lb_replenish, sizeof(int64_t));
  get("UBREPLENISHflagTOutput", &S.ub_replenish, sizeof(int64_t));
  if (!bad.empty()) { g_init_err = "cannot read " + bad + " (missing or not 8 bytes)"; return KAOS_E_IO; }
  S.nq = S.nq_ub - S.nq_lb + 1;
  if (S.nq < 1 || S.nq > (1LL << 30)) {
    g_init_err = "spin-up files give NqUBT - NqLBT + 1 = " + std::to_string(S.nq) + " cells";
    return KAOS_E_IO;
  }
  if (S.nq > nq_cap) {
    g_init_err = "spin-up files hold " + std::to_string(S.nq) + " cells, more than nq_cap";
    return KAOS_E_ARG;
  }
  const size_t nb = sizeof(double) * (size_t)S.nq;
  std::vector<double> tmp(S.nq);
  const char* names[5] = {"DensityOutputRT", "TemperatureOutputRT", "EAInertialOutputRT", "EAPressureOutputRT",
                          "EAmagOutputRT"};
  double* outs[5] = {density, temperature, ea_inertial, ea_pressure, ea_mag};
  for (int i = 0; i < 5; ++i) {
    get(names[i], tmp.data(), nb);         // every file is read (and checked) even if not wanted
    if (bad.empty() && outs[i]) memcpy(outs[i], tmp.data(), nb);
  }
  if (!bad.empty()) {
    g_init_err = "cannot read " + bad + " (missing or not " + std::to_string(S.nq) + " binary64 values)";
    return KAOS_E_IO;
  }
  *st = S;
  return KAOS_OK;
}

extern "C" int kaos_set_field(kaos_ctx* c, int32_t tube, const double* EA, int32_t mode) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->ht[tube].set) return fail(c, KAOS_E_STATE, "kaos_set_field before kaos_set_grid");
  if (mode != KAOS_FIELD_COMPUTED && mode != KAOS_FIELD_INPUT_FIRST && mode != KAOS_FIELD_INPUT_FIXED)
    return fail(c, KAOS_E_ARG, "field mode must be KAOS_FIELD_COMPUTED, _INPUT_FIRST or _INPUT_FIXED");
  if (mode != KAOS_FIELD_COMPUTED && !EA) return fail(c, KAOS_E_ARG, "an input field needs EA");
  HostTube& h = c->ht[tube];
  const int nq = h.g.nq;
  if (EA) {
    for (int Q = 0; Q < nq; ++Q)
      if (!std::isfinite(EA[Q])) return fail(c, KAOS_E_ARG, "input field value not finite");
    h.ea_in.assign(EA, EA + nq);
  }
  h.ea_mode = mode;
  if (c->built) {
    // device copies: the tube table (graphs read it at replay) and the values
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevTube& T = c->dt[tube];
    T.ea_in = mode;
    HIPCHK(c, hipMemcpy(c->d_tubes + tube, &T, sizeof(DevTube), hipMemcpyHostToDevice));
    if (EA) HIPCHK(c, hipMemcpy(c->d_EAin + T.goff, h.ea_in.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
  }
  return KAOS_OK;
}

// Charge-exchange contexts keep ENAs and collision counts that the handover
// and restart files do not hold.
static bool any_cx_tube(const kaos_ctx* c) {
  for (const HostTube& h : c->ht) if (h.set && h.g.cx) return true;
  return false;
}

extern "C" int kaos_download_cx(kaos_ctx* c, int32_t tube, double* sigma, double* nu, double* poi) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built || !c->dt[tube].cx) return fail(c, KAOS_E_STATE, "no charge exchange in this tube");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const size_t nb = sizeof(double) * T.nq;
  if (sigma) HIPCHK(c, hipMemcpy(sigma, c->d_cx + 3 * T.goff, nb, hipMemcpyDeviceToHost));
  if (nu) HIPCHK(c, hipMemcpy(nu, c->d_cx + 3 * T.goff + T.nq, nb, hipMemcpyDeviceToHost));
  if (poi) HIPCHK(c, hipMemcpy(poi, c->d_cx + 3 * T.goff + 2 * T.nq, nb, hipMemcpyDeviceToHost));
  return KAOS_OK;
}

extern "C" int kaos_download_enas(kaos_ctx* c, int32_t tube, int64_t n_cap, int64_t* n_out, double* x, double* y,
                                  double* z, double* Vx, double* Vy, double* Vz, int64_t* id, int64_t* qind) {
  if (!c || !n_out) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  *n_out = 0;
  if (!c->built) return KAOS_OK;
  const DevTube& T = c->dt[tube];
  if (!T.cx) return KAOS_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  long long used = 0;
  HIPCHK(c, hipMemcpy(&used, c->d_ena_cnt + tube, sizeof(long long), hipMemcpyDeviceToHost));
  used = std::min(used, T.ecap);
  std::vector<double> X(6 * (size_t)std::max(used, 1LL));
  std::vector<long long> ids(std::max(used, 1LL));
  std::vector<int> q(std::max(used, 1LL));
  for (int i = 0; i < 6 && used; ++i)
    HIPCHK(c, hipMemcpy(X.data() + (size_t)i * used, c->d_ena + (size_t)i * c->ena_tot + T.eoff,
                        sizeof(double) * used, hipMemcpyDeviceToHost));
  if (used) {
    HIPCHK(c, hipMemcpy(ids.data(), c->d_eid + T.eoff, sizeof(long long) * used, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(q.data(), c->d_eq + T.eoff, sizeof(int) * used, hipMemcpyDeviceToHost));
  }
  std::vector<long long> ord;
  for (long long e = 0; e < used; ++e) if (q[e] != KAOS_ENA_DEAD) ord.push_back(e);
  std::sort(ord.begin(), ord.end(), [&](long long u, long long v) { return ids[u] < ids[v]; });
  const long long n = (long long)ord.size();
  if (n > n_cap) return fail(c, KAOS_E_ARG, "n_cap smaller than the tube's " + std::to_string(n) + " ENAs");
  double* out[6] = {x, y, z, Vx, Vy, Vz};
  for (long long i = 0; i < n; ++i) {
    const long long e = ord[i];
    for (int k = 0; k < 6; ++k) if (out[k]) out[k][i] = X[(size_t)k * used + e];
    if (id) id[i] = ids[e];
    if (qind) qind[i] = q[e];
  }
  *n_out = n;
  return KAOS_OK;
}

extern "C" int kaos_download_ena_moments(kaos_ctx* c, int32_t tube, int64_t* counts, double* M) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built || !c->dt[tube].cx) return fail(c, KAOS_E_STATE, "no charge exchange in this tube");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  if (counts) {
    const size_t n = (size_t)T.nq * T.nvp * T.nvq * T.nvphi;
    std::vector<unsigned int> u(n);
    HIPCHK(c, hipMemcpy(u.data(), c->d_ena_hist + T.ehoff, sizeof(unsigned int) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) counts[i] = u[i];
  }
  if (M) HIPCHK(c, hipMemcpy(M, c->d_ena_M + 8 * T.goff, sizeof(double) * 8 * T.nq, hipMemcpyDeviceToHost));
  return KAOS_OK;
}

extern "C" int kaos_download_ena_master(kaos_ctx* c, int32_t tube, int64_t* nq_ena, double* esc) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built || !c->dt[tube].cx) return fail(c, KAOS_E_STATE, "no charge exchange in this tube");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const HostTube& h = c->ht[tube];
  if (nq_ena) {
    std::vector<unsigned int> u(T.nq);
    HIPCHK(c, hipMemcpy(u.data(), c->d_nq_ena + T.goff, sizeof(unsigned int) * T.nq, hipMemcpyDeviceToHost));
    for (int Q = 0; Q < T.nq; ++Q) nq_ena[Q] = u[Q];
  }
  if (esc) {
    unsigned long long e2[2];
    HIPCHK(c, hipMemcpy(e2, c->d_ena_esc + 2 * tube, sizeof(e2), hipMemcpyDeviceToHost));
    // fluxes as BoundaryConditionsSub writes them (BC:355-366): count nsnormfac over
    // (d3x / (hq dq)) of the boundary cell times ndatfac h
    const int ub = T.nq - 1;
    const double dtm = (double)c->cfg.ndatfac * T.h;
    esc[0] = (double)e2[0]; esc[1] = (double)e2[1];
    esc[2] = ((double)e2[0] * T.nsnormfac) / ((h.d3x[0] / (h.hq[0] * h.dq[0])) * dtm);
    esc[3] = ((double)e2[1] * T.nsnormfac) / ((h.d3x[ub] / (h.hq[ub] * h.dq[ub])) * dtm);
  }
  return KAOS_OK;
}

// The live ions of a tube on the host, in the reference's order (survivors
// in order, then injected ions: by id), with their Cartesian state.
struct TubeDump {
  long long n = 0;                              // slots
  std::vector<double> xyz, vr, vr2, qk, ae, ag;
  std::vector<uint32_t> cell;
  std::vector<long long> ids, order;            // order: live slots sorted by id
};

static int fetch_tube(kaos_ctx* c, int32_t tube, TubeDump& d) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevTube& T = c->dt[tube];
  const long long n = c->ht[tube].count;
  const bool mixed = c->cfg.precision == KAOS_MIXED;
  d.n = n;
  d.xyz.assign(6 * (size_t)n, 0.0);
  d.vr.assign(n, 0.0); d.vr2.assign(n, 0.0); d.qk.assign(n, 0.0); d.ae.assign(n, 0.0); d.ag.assign(n, 0.0);
  d.cell.assign(n, 0u); d.ids.assign(n, 0); d.order.clear();
  if (!n) return KAOS_OK;
  std::vector<char> live(n, 0);
  double* d6 = nullptr;
  HIPCHK(c, hipMalloc((void**)&d6, sizeof(double) * 6 * n));
  StepArgs a = make_args(c, 0, 0, 0, 0);
  hipLaunchKernelGGL(k_reconstruct, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, T, n, a,
                     (int)mixed, d6);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(d.xyz.data(), d6, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
  (void)hipFree(d6);
  const StateBufs& B = c->st[c->cur];
  HIPCHK(c, hipMemcpy(d.cell.data(), B.cell + T.off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(d.ids.data(), B.id + T.off, sizeof(long long) * n, hipMemcpyDeviceToHost));
  if (mixed) {
    std::vector<float> f(n);
    float* src[5] = {B.vrf, B.vrf2, B.qkf, B.aef, B.agf};
    std::vector<double>* dst[5] = {&d.vr, &d.vr2, &d.qk, &d.ae, &d.ag};
    for (int i = 0; i < 5; ++i) {
      HIPCHK(c, hipMemcpy(f.data(), src[i] + T.off, sizeof(float) * n, hipMemcpyDeviceToHost));
      for (long long j = 0; j < n; ++j) (*dst[i])[j] = (double)f[j];
    }
    const double* qGC = c->ht[tube].qGC.data();
    for (long long j = 0; j < n; ++j) {
      live[j] = d.cell[j] != KAOS_DEAD;
      if (live[j]) d.qk[j] += kx_qref(qGC[kx_refq((int)((d.cell[j] >> 16) & KAOS_QMASK) - 1, T.nq) - 1]);
    }
  } else {
    std::vector<double> rs(n);
    HIPCHK(c, hipMemcpy(rs.data(), B.rs + T.off, sizeof(double) * n, hipMemcpyDeviceToHost));
    double* src[5] = {B.vr, B.vr2, B.qk, B.ae, B.ag};
    std::vector<double>* dst[5] = {&d.vr, &d.vr2, &d.qk, &d.ae, &d.ag};
    for (int i = 0; i < 5; ++i)
      HIPCHK(c, hipMemcpy(dst[i]->data(), src[i] + T.off, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (long long j = 0; j < n; ++j) live[j] = rs[j] != 0.0;
  }
  d.order.reserve(n);
  for (long long j = 0; j < n; ++j) if (live[j]) d.order.push_back(j);
  std::sort(d.order.begin(), d.order.end(), [&](long long p, long long q) { return d.ids[p] < d.ids[q]; });
  return KAOS_OK;
}

extern "C" int kaos_download_particles(kaos_ctx* c, int32_t tube, int64_t n_cap, int64_t* n_out,
                                       double* x, double* y, double* z, double* Vx, double* Vy, double* Vz,
                                       double* vperp, double* qk4, double* aea, double* ag,
                                       int64_t* qind, int64_t* vparind, int64_t* vperpind) {
  if (!c || !n_out) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built) { *n_out = 0; return KAOS_OK; }
  TubeDump d;
  int rc = fetch_tube(c, tube, d);
  if (rc) return rc;
  if ((long long)d.order.size() > n_cap) return fail(c, KAOS_E_CAPACITY, "output arrays too small");
  long long k = 0;
  for (long long j : d.order) {
    if (x) x[k] = d.xyz[6 * j + 0];
    if (y) y[k] = d.xyz[6 * j + 1];
    if (z) z[k] = d.xyz[6 * j + 2];
    if (Vx) Vx[k] = d.xyz[6 * j + 3];
    if (Vy) Vy[k] = d.xyz[6 * j + 4];
    if (Vz) Vz[k] = d.xyz[6 * j + 5];
    if (vperp) vperp[k] = fabs(sqrt(d.vr[j] * d.vr[j] + d.vr2[j] * d.vr2[j]));   // |(Vperp1, Vperp2)|
    if (qk4) qk4[k] = d.qk[j];
    if (aea) aea[k] = d.ae[j];
    if (ag) ag[k] = d.ag[j];
    if (qind) qind[k] = (int64_t)((d.cell[j] >> 16) & KAOS_QMASK) - 1;
    if (vperpind) vperpind[k] = (int64_t)((d.cell[j] >> 8) & 0xFFu);
    if (vparind) vparind[k] = (int64_t)(d.cell[j] & 0xFFu);
    ++k;
  }
  *n_out = k;
  return KAOS_OK;
}

extern "C" int kaos_download_vperp(kaos_ctx* c, int32_t tube, int64_t n_cap, int64_t* n_out, double* vperp1,
                                   double* vperp2) {
  if (!c || !n_out) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!c->built) { *n_out = 0; return KAOS_OK; }
  TubeDump d;
  int rc = fetch_tube(c, tube, d);
  if (rc) return rc;
  if ((long long)d.order.size() > n_cap) return fail(c, KAOS_E_CAPACITY, "output arrays too small");
  long long k = 0;
  for (long long j : d.order) {
    if (vperp1) vperp1[k] = d.vr[j];
    if (vperp2) vperp2[k] = d.vr2[j];
    ++k;
  }
  *n_out = k;
  return KAOS_OK;
}

extern "C" int kaos_upload_state(kaos_ctx* c, int32_t tube, int64_t n, const double* x, const double* y,
                                 const double* z, const double* Vx, const double* Vy, const double* Vz,
                                 const double* vperp1, const double* vperp2, const double* vperp,
                                 const double* aea, const double* ag, const double* qk4, const int64_t* qind,
                                 const int64_t* vparind, const int64_t* vperpind) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (any_cx_tube(c)) return fail(c, KAOS_E_STATE, "kaos_upload_state: not with charge exchange (ENAs)");
  if (n < 0) return fail(c, KAOS_E_ARG, "negative ion count");
  if (n > 0 && (!x || !y || !z || !Vx || !Vy || !Vz || !vperp || !aea || !ag || !qk4 || !qind || !vparind ||
                !vperpind))
    return fail(c, KAOS_E_ARG, "bad particle arrays (only vperp1 and vperp2 may be NULL)");
  if ((vperp1 == nullptr) != (vperp2 == nullptr)) return fail(c, KAOS_E_ARG, "vperp1 and vperp2 go together");
  if (!c->built) { int rc = build_device(c); if (rc) return rc; }
  HostTube& h = c->ht[tube];
  const DevTube& T = c->dt[tube];
  if (n > h.cap) return fail(c, KAOS_E_CAPACITY, "more ions than the tube capacity");
  for (int64_t i = 0; i < n; ++i) {
    if (qind[i] < -1 || qind[i] > T.nq || vparind[i] < 0 || vparind[i] > T.nvpar || vperpind[i] < 0 ||
        vperpind[i] > T.nvperp)
      return fail(c, KAOS_E_ARG, "cell index out of range at ion " + std::to_string(i));
    if (!(x[i] != 0.0 || y[i] != 0.0 || z[i] != 0.0) || !std::isfinite(x[i] + y[i] + z[i]))
      return fail(c, KAOS_E_ARG, "position not finite or at the origin at ion " + std::to_string(i));
  }
  if (n) {
    const double* src[12] = {x, y, z, Vx, Vy, Vz, vperp1, vperp2, vperp, aea, ag, qk4};
    const int64_t* isrc[3] = {qind, vparind, vperpind};
    double* dd[12] = {};
    long long* di[3] = {};
    int rc = KAOS_OK;
    for (int k = 0; k < 12 && !rc; ++k) {
      if (!src[k]) continue;
      if (hipMalloc((void**)&dd[k], sizeof(double) * n) != hipSuccess ||
          hipMemcpy(dd[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, KAOS_E_DEVICE, "kaos_upload_state: device staging failed");
    }
    for (int k = 0; k < 3 && !rc; ++k)
      if (hipMalloc((void**)&di[k], sizeof(long long) * n) != hipSuccess ||
          hipMemcpy(di[k], isrc[k], sizeof(long long) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, KAOS_E_DEVICE, "kaos_upload_state: device staging failed");
    if (!rc) {
      UpState u = {dd[0], dd[1], dd[2], dd[3], dd[4], dd[5], dd[6], dd[7], dd[8], dd[9], dd[10], dd[11],
                   di[0], di[1], di[2]};
      StepArgs a = make_args(c, 0, 0, 0, 0);
      hipLaunchKernelGGL(k_upload_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, T,
                         (long long)n, u, 0LL, c->d_vparGH + T.voff, c->d_vperpGH + T.woff, c->d_rref + T.goff, a,
                         (int)(c->cfg.precision == KAOS_MIXED));
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(c, KAOS_E_DEVICE, "kaos_upload_state: kernel failed");
    }
    for (double* p : dd) if (p) (void)hipFree(p);
    for (long long* p : di) if (p) (void)hipFree(p);
    if (rc) return rc;
  }
  // slots past the new count are dead (a mid-run upload may replace more ions)
  if (h.count > n) {
    const long long nd = h.count - n;
    const StateBufs& B = c->st[c->cur];
    HIPCHK(c, hipMemsetAsync(B.cell + T.off + n, 0xFF, sizeof(uint32_t) * nd, c->stream));
    if (B.rs) HIPCHK(c, hipMemsetAsync(B.rs + T.off + n, 0, sizeof(double) * nd, c->stream));
  }
  long long cnt = n, nid = 0;
  HIPCHK(c, hipMemcpy(&nid, c->d_next_id + tube, sizeof(long long), hipMemcpyDeviceToHost));
  if (nid < cnt) nid = cnt;
  HIPCHK(c, hipMemcpyAsync(c->d_count + tube, &cnt, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_live + tube, &cnt, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_next_id + tube, &nid, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  h.count = n;
  h.uploaded = true;
  return KAOS_OK;
}

extern "C" int kaos_upload_moments(kaos_ctx* c, int32_t tube, const double* M) {
  if (!c) return KAOS_E_ARG;
  if (tube < 0 || tube >= c->cfg.ntubes) return fail(c, KAOS_E_ARG, "tube out of range");
  if (!M) return fail(c, KAOS_E_ARG, "no moments");
  if (!c->built) { int rc = build_device(c); if (rc) return rc; }
  const DevTube& T = c->dt[tube];
  HIPCHK(c, hipMemcpy(c->d_M + 6 * T.goff, M, sizeof(double) * 6 * T.nq, hipMemcpyHostToDevice));
  return KAOS_OK;
}

extern "C" int kaos_resume(kaos_ctx* c, int64_t n_last) {
  if (!c) return KAOS_E_ARG;
  if (n_last < 1) return fail(c, KAOS_E_ARG, "n_last must be >= 1");
  if (c->last_step != 0) return fail(c, KAOS_E_STATE, "kaos_resume needs a context that has not stepped");
  if (!c->built) return fail(c, KAOS_E_STATE, "kaos_resume needs the tubes' ions (kaos_upload_state)");
  for (int t = 0; t < c->cfg.ntubes; ++t)
    if (!c->ht[t].uploaded) return fail(c, KAOS_E_STATE, "tube " + std::to_string(t) + " has no ions");
  long long s = n_last;
  HIPCHK(c, hipMemcpy(c->d_stepno, &s, sizeof(long long), hipMemcpyHostToDevice));
  c->last_step = n_last;
  c->last_master = (n_last == 1) ? 1 : (int)is_master(c, n_last);
  return KAOS_OK;
}

// Checkpoint file: a header, the context's step state and, per tube, its
// bookkeeping and the slots [0, count) of every state array of the live
// buffer, as stored (so the continuation is bit for bit the uninterrupted run
// in either precision form).
static const char KAOS_CKPT_MAGIC[8] = {'K', 'A', 'O', 'S', 'C', 'K', 'P', '2'};

static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
  return h;
}

// Hash of everything that fixes a context's layout and arithmetic: the run
// configuration and every tube's grid, species and flags.
static uint64_t ctx_fingerprint(const kaos_ctx* c) {
  uint64_t h = 1469598103934665603ull;
  const int32_t cf[4] = {c->cfg.ntubes, c->cfg.precision, c->cfg.ndatfac, c->cfg.flags & KAOS_FLAG_VPERP3D};
  h = fnv1a(h, cf, sizeof(cf));
  h = fnv1a(h, &c->cfg.seed, sizeof(c->cfg.seed));
  for (const HostTube& t : c->ht) {
    const kaos_grid& g = t.g;
    const int32_t gi[14] = {g.nq, g.nvpar, g.nvperp, g.grav, g.eamb, g.eambsign, g.filter_pt, g.lbcond,
                            g.ubcond, g.filter_pt_m1par, g.icr, g.nvperp1, g.nvperp2, 0};
    h = fnv1a(h, gi, sizeof(gi));
    const double gd[10] = {g.p, g.cphi, g.sphi, g.h, g.ms, g.qs, g.nsnormfac, g.Ts_lb, g.Ts_ub,
                           (double)(g.n_inj_lb + 3 * g.n_inj_ub)};
    h = fnv1a(h, gd, sizeof(gd));
    const std::vector<double>* arr[] = {&t.qGL, &t.qGH, &t.qGC, &t.hq, &t.dq, &t.d3x, &t.rGC, &t.cth, &t.ell,
                                        &t.Te, &t.vparGL, &t.vparGH, &t.vperpGL, &t.vperpGH, &t.d3v,
                                        &t.v1GL, &t.v1GH, &t.v2GL, &t.v2GH};
    for (const std::vector<double>* v : arr) h = fnv1a(h, v->data(), sizeof(double) * v->size());
  }
  return h;
}

// The device arrays of a checkpoint: (pointer, element size, elements) per tube
// slot array, then the per-cell and per-tube context arrays.
static std::vector<std::pair<void*, size_t>> slot_arrays(kaos_ctx* c) {
  StateBufs& B = c->st[c->cur];
  std::vector<std::pair<void*, size_t>> v;
  if (c->cfg.precision == KAOS_MIXED) {
    v = {{B.rdi, 4}, {B.vpi, 4}, {B.lo, 4}, {B.vrf, 4}, {B.vrf2, 4}, {B.qkf, 4}, {B.aef, 4}, {B.agf, 4}};
    if (B.c12) v.push_back({B.c12, 2});
  } else {
    v = {{B.rs, 8}, {B.vp, 8}, {B.vr, 8}, {B.vr2, 8}, {B.qk, 8}, {B.ae, 8}, {B.ag, 8}};
  }
  v.push_back({B.cell, 4});
  v.push_back({B.id, 8});
  return v;
}

static std::vector<std::pair<void*, size_t>> ctx_arrays(kaos_ctx* c) {
  const size_t g = (size_t)c->ncell_tot, nt = (size_t)c->cfg.ntubes;
  std::vector<std::pair<void*, size_t>> v = {
      {c->d_M, 8 * 6 * g}, {c->d_M0F, 8 * g}, {c->d_M1F, 8 * g}, {c->d_EA, 8 * g}, {c->d_NqRT, 8 * g},
      {c->d_nqc, 4 * g}, {c->d_nqoff, 4 * g}, {c->d_hist, 4 * (size_t)c->nhist_tot},
      {c->d_esc_acc, 8 * 2 * nt}, {c->d_esc_rep, 8 * 2 * nt}, {c->d_stepno, 8}, {c->d_pushed, 8},
      {c->d_count, 8 * nt}, {c->d_live, 8 * nt}, {c->d_next_id, 8 * nt}};
  if (c->v3d) {
    v.push_back({c->d_M3, 8 * 8 * g});
    v.push_back({c->d_hist3, 4 * (size_t)c->nhist3_tot});
  }
  return v;
}

extern "C" int kaos_save_state(kaos_ctx* c, const char* path) {
  if (!c || !path) return KAOS_E_ARG;
  if (!c->built || c->last_step < 1) return fail(c, KAOS_E_STATE, "kaos_save_state needs a context that has stepped");
  if (any_cx_tube(c)) return fail(c, KAOS_E_STATE, "kaos_save_state: not with charge exchange (ENAs)");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  FILE* f = fopen(path, "wb");
  if (!f) return fail(c, KAOS_E_IO, std::string("cannot write ") + path);
  bool ok = true;
  auto put = [&](const void* p, size_t n) { if (ok && n) ok = fwrite(p, 1, n, f) == n; };
  const uint64_t fp = ctx_fingerprint(c);
  const long long hdr[4] = {c->last_step, c->last_master, c->masters_since_reorder, c->cfg.ntubes};
  put(KAOS_CKPT_MAGIC, 8);
  put(&fp, 8);
  put(hdr, sizeof(hdr));
  std::vector<char> buf;
  int rc = KAOS_OK;
  for (auto& a : ctx_arrays(c)) {
    buf.resize(a.second);
    if (hipMemcpy(buf.data(), a.first, a.second, hipMemcpyDeviceToHost) != hipSuccess) { rc = KAOS_E_DEVICE; break; }
    put(buf.data(), a.second);
  }
  for (int t = 0; t < c->cfg.ntubes && !rc; ++t) {
    const long long n = c->ht[t].count, off = c->dt[t].off;
    put(&n, 8);
    for (auto& a : slot_arrays(c)) {
      buf.resize(a.second * (size_t)n);
      if (n && hipMemcpy(buf.data(), (char*)a.first + a.second * off, a.second * n, hipMemcpyDeviceToHost) !=
                   hipSuccess) { rc = KAOS_E_DEVICE; break; }
      put(buf.data(), a.second * (size_t)n);
    }
  }
  ok = (fclose(f) == 0) && ok;
  if (rc) return fail(c, rc, "kaos_save_state: device copy failed");
  if (!ok) return fail(c, KAOS_E_IO, std::string("short write to ") + path);
  return KAOS_OK;
}

extern "C" int kaos_load_state(kaos_ctx* c, const char* path) {
  if (!c || !path) return KAOS_E_ARG;
  if (c->last_step != 0) return fail(c, KAOS_E_STATE, "kaos_load_state needs a context that has not stepped");
  if (any_cx_tube(c)) return fail(c, KAOS_E_STATE, "kaos_load_state: not with charge exchange (ENAs)");
  if (!c->built) { int rc = build_device(c); if (rc) return rc; }
  FILE* f = fopen(path, "rb");
  if (!f) return fail(c, KAOS_E_IO, std::string("cannot read ") + path);
  bool ok = true;
  auto get = [&](void* p, size_t n) { if (ok && n) ok = fread(p, 1, n, f) == n; };
  char magic[8] = {};
  uint64_t fp = 0;
  long long hdr[4] = {};
  get(magic, 8); get(&fp, 8); get(hdr, sizeof(hdr));
  if (!ok || memcmp(magic, KAOS_CKPT_MAGIC, 8) != 0) { fclose(f); return fail(c, KAOS_E_ARG, "not a kaos checkpoint"); }
  if (fp != ctx_fingerprint(c) || hdr[3] != c->cfg.ntubes) {
    fclose(f);
    return fail(c, KAOS_E_ARG, "checkpoint of another configuration (kaos_cfg or a tube's grid differs)");
  }
  std::vector<char> buf;
  int rc = KAOS_OK;
  for (auto& a : ctx_arrays(c)) {
    buf.resize(a.second);
    get(buf.data(), a.second);
    if (!ok) break;
    if (hipMemcpy(a.first, buf.data(), a.second, hipMemcpyHostToDevice) != hipSuccess) { rc = KAOS_E_DEVICE; break; }
  }
  std::vector<long long> cnt(c->cfg.ntubes, 0);
  for (int t = 0; t < c->cfg.ntubes && ok && !rc; ++t) {
    long long n = 0;
    get(&n, 8);
    if (!ok) break;
    if (n < 0 || n > c->ht[t].cap) { fclose(f); return fail(c, KAOS_E_CAPACITY, "checkpoint holds more slots than the tube capacity"); }
    cnt[t] = n;
    const long long off = c->dt[t].off;
    for (auto& a : slot_arrays(c)) {
      buf.resize(a.second * (size_t)n);
      get(buf.data(), a.second * (size_t)n);
      if (!ok) break;
      if (n && hipMemcpy((char*)a.first + a.second * off, buf.data(), a.second * n, hipMemcpyHostToDevice) !=
                   hipSuccess) { rc = KAOS_E_DEVICE; break; }
    }
  }
  fclose(f);
  if (rc) return fail(c, rc, "kaos_load_state: device copy failed");
  if (!ok) return fail(c, KAOS_E_IO, std::string("truncated checkpoint ") + path);
  for (int t = 0; t < c->cfg.ntubes; ++t) { c->ht[t].count = cnt[t]; c->ht[t].uploaded = true; }
  c->last_step = hdr[0];
  c->last_master = (int)hdr[1];
  c->masters_since_reorder = (int)hdr[2];
  return KAOS_OK;
}

extern "C" int kaos_set_launch(kaos_ctx* c, int32_t tiles_per_group, int32_t hist_words, int32_t persistent) {
  if (!c) return KAOS_E_ARG;
  if (tiles_per_group < 1 || tiles_per_group * KAOS_TILE > 65535)
    return fail(c, KAOS_E_ARG, "tiles_per_group must be in [1, 63] (16-bit LDS counters)");
  if (hist_words < 0 || hist_words > 12288) return fail(c, KAOS_E_ARG, "hist_words must be in [0, 12288]");
  c->launch = {tiles_per_group, hist_words, persistent ? 1 : 0, c->launch.block};
  c->tuning = false;
  for (auto& ev : c->tune_ev) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
  c->tune_ev.clear();
  return KAOS_OK;
}

extern "C" int kaos_get_launch(kaos_ctx* c, int32_t* tiles_per_group, int32_t* hist_words, int32_t* persistent,
                               int32_t* tuning) {
  if (!c) return KAOS_E_ARG;
  if (tiles_per_group) *tiles_per_group = c->launch.tpb;
  if (hist_words) *hist_words = c->launch.hwords;
  if (persistent) *persistent = c->launch.persistent;
  if (tuning) *tuning = c->tuning ? 1 : 0;
  return KAOS_OK;
}

extern "C" int kaos_set_launch_block(kaos_ctx* c, int32_t block) {
  if (!c) return KAOS_E_ARG;
  if (block != 256 && block != 512) return fail(c, KAOS_E_ARG, "push block must be 256 or 512 threads");
  c->launch.block = block;
  c->tuning = false;
  for (auto& ev : c->tune_ev) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
  c->tune_ev.clear();
  return KAOS_OK;
}

extern "C" int kaos_get_launch_block(kaos_ctx* c, int32_t* block) {
  if (!c || !block) return KAOS_E_ARG;
  *block = (c->launch.block == 512 && (!c->built || plain_push(c))) ? 512 : 256;
  return KAOS_OK;
}

extern "C" int kaos_profile(kaos_ctx* c, int32_t enable) {
  if (!c) return KAOS_E_ARG;
  c->prof_on = enable != 0;
  return KAOS_OK;
}

extern "C" int kaos_profile_read(kaos_ctx* c, double* push_ms, int64_t* push_launches, double* run_ms) {
  if (!c) return KAOS_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tp = 0.0, tr = 0.0;
  for (auto& e : c->ev_push) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e.first, e.second));
    tp += ms;
    (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second);
  }
  for (auto& e : c->ev_step) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e.first, e.second));
    tr += ms;
    (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second);
  }
  if (push_ms) *push_ms = tp;
  if (push_launches) *push_launches = (int64_t)c->ev_push.size();
  if (run_ms) *run_ms = tr;
  c->ev_push.clear(); c->ev_step.clear();
  return KAOS_OK;
}

extern "C" int kaos_finalize(kaos_ctx* c) {
  if (!c) return KAOS_OK;
  (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->d_tubes, c->d_tile_tube, c->d_tile_start, c->d_blk_tube, c->d_blk_q, c->d_cells_wave,
                  c->d_cells_block, c->d_qGH, c->d_qGC,
                  c->d_hq, c->d_dq, c->d_Te, c->d_d3x, c->d_EA, c->d_EAin, c->d_EG, c->d_M, c->d_M0F, c->d_M1F, c->d_NqRT,
                  c->d_vparGH, c->d_vperpGH, c->d_vparGC, c->d_vperpGC, c->d_dvpar, c->d_dvperp, c->d_d3v,
                  c->d_rGH, c->d_rref, c->d_rk, c->d_vk, c->d_keycnt, c->d_cursor, c->d_cnt_new,
                  c->d_next_id, c->d_count, c->d_hist,
                  c->d_nqc, c->d_nqoff, c->d_status, c->d_esc_acc, c->d_esc_rep, c->d_inj_base, c->d_inj_cnt,
                  c->d_uni_base, c->d_uni, c->d_pushed};
  for (void* p : bufs) if (p) (void)hipFree(p);
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& ev : c->tune_ev) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
  (void)hipFree(c->d_work);
  (void)hipFree(c->d_live);
  (void)hipFree(c->d_vbox); (void)hipFree(c->d_hbox);
  (void)hipFree(c->d_vmarg_par); (void)hipFree(c->d_vmarg_perp);
  (void)hipFree(c->d_stepno);
  void* b3[] = {c->d_v1GH, c->d_v1GC, c->d_dv1, c->d_v2GH, c->d_v2GC, c->d_dv2, c->d_d3v3, c->d_M3, c->d_hist3,
                c->d_grp3};
  for (void* p : b3) if (p) (void)hipFree(p);
  void* bcx[] = {c->d_nsneut, c->d_cxgeo, c->d_cx, c->d_cx_need, c->d_cx_pref, c->d_cx_thr, c->d_cx_hist, c->d_cx_lev,
                 c->d_ena, c->d_eid, c->d_ena_cnt, c->d_eoff, c->d_eq, c->d_nq_ena, c->d_ena_esc,
                 c->d_ena_vp, c->d_ena_vq, c->d_ena_vf, c->d_ena_hh, c->d_ena_d33v, c->d_ena_M, c->d_evc,
                 c->d_ena_hist};
  for (void* p : bcx) if (p) (void)hipFree(p);
  for (StateBufs& B : c->st) {
    void* sb[] = {B.rs, B.vp, B.vr, B.qk, B.ae, B.ag, B.rdi, B.vpi, B.lo, B.vrf, B.qkf, B.aef, B.agf, B.cell, B.id,
                  B.vr2, B.vrf2, B.c12, B.ck4, B.ck4f};
    for (void* p : sb) if (p) (void)hipFree(p);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KAOS_OK;
}
