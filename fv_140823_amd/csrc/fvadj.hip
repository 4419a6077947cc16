/*
 * fvadj.hip — discrete adjoint and tangent-linear model of RHS + RK rollouts.
 *
 * Objective of a rollout of nsteps SSP-RK3 steps at a fixed dt (RK.c:24-47):
 *     J = sum_{n=1..nsteps} dt * sum_{i in probe} p(S^n_i) Vol_i
 * Adjoint (reverse, hand written): for each stage
 *     S_k = a SN + b S_{k-1} + c dt/Vol R(S_{k-1})
 * lambda_SN += a lambda, lambda_{k-1} = b lambda + R'(S_{k-1})^T (c dt/Vol lambda),
 * with R'^T = recon^T . face^T . gather^T, each a deterministic gather:
 *   k_adj_face        per face: MUSCL face states (the forward's own
 *                     reconstruction), mu = c dt lam / Vol on both sides,
 *                     mu_L - mu_R -> hand-written VJP of HLLC + EOS
 *   k_adj_recon       per cell: face-state bars by gather slot (f = 0..5, k = 0..3),
 *                     limiter partials -> own bar and the bars sent to the +/- neighbours;
 *                     over 512-cell blocks, a bar for a same-level neighbour of the block
 *                     handed over in LDS, the others written out
 *   k_adj_recv_stage  per cell: own bar + what its other neighbours sent (no atomics),
 *                     then the stage's reverse update of lam and lamSN
 * Stage states are recomputed from checkpoints every `ckpt` steps, written
 * straight into a segment store that the reverse sweep reads in place.
 * Tangent (forward, for the dot-product check): the same stage kernels on
 * forward-mode dual numbers for the face flux and analytic limiter partials.
 * Navier-Stokes: the face VJP also returns the bars of both cells' momentum
 * gradients (visc_vjp_ideal), k_adj_sg gathers them per cell, and
 * k_adj_recv_stage adds the transpose of the Green-Gauss gradient (gradcalc.c:
 * 23-42): the cell's own term and its face neighbours' terms.
 * Scope: model 1 (ideal/stiffened gas), Euler or Navier-Stokes, MUSCL (either
 * limiter) or first order; bc outflow/periodic (an inlet ghost is a constant).
 */
#include <cmath>

#include "fv_common.h"

struct fv_adj_state {
  int nsteps = 0, ckpt = 0, nck = 0;
  double dt = 0.0;
  bool fwd_done = false;
  size_t plane = 0; /* neq * ns doubles */
  int seg_valid = -1; /* segment whose stage states the store holds, or -1 */
  uint8_t *mask = nullptr;
  double *ckpts = nullptr, *seg = nullptr;
  double *lam = nullptr, *lamSN = nullptr, *self = nullptr;
  double *Pp = nullptr, *Pm = nullptr;
  void *lamUL = nullptr, *lamUR = nullptr;
  void *lamG = nullptr;      /* [nfp][9] gradient bars per face side (Navier-Stokes) */
  double *SGb = nullptr;     /* [9][ns] gradient bars per cell */
  double *SGmd = nullptr;    /* [9][ns] tangent of the momentum gradients */
  double *Jsteps = nullptr, *Jdsteps = nullptr, *part = nullptr; /* (hi, lo) pairs */
  double J_hi = 0.0, J_lo = 0.0; /* objective of the last rollout, unevaluated sum */
  int part_cap = 0;
  double *Sd = nullptr, *SNd = nullptr, *SFd = nullptr, *Gd = nullptr;
};

void fv_adj_release(fvg_ctx *c) {
  if (!c || !c->adj) return;
  fv_adj_state *a = c->adj;
  void *ptrs[] = {a->mask, a->ckpts, a->seg, a->lam, a->lamSN, a->self, a->Pp, a->Pm, a->lamUL,
                  a->lamUR, a->lamG, a->SGb, a->SGmd, a->Jsteps, a->Jdsteps, a->part, a->Sd, a->SNd,
                  a->SFd, a->Gd};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete a;
  c->adj = nullptr;
}

/* ------------------------------------------------------------------ kernels */

__device__ __forceinline__ double cell_vol(const DevMesh &M, int i) {
  return M.vol[(M.cflag[i] >> 16) & 0xff];
}

/* face^T: lamUL/lamUR[j][v] = d(sum_i mu_i . R_i)/d(face states of face j), with
 * mu_i = cdt lam_i / Vol_i.  The face states are rebuilt exactly as the forward
 * face kernel builds them (recon_face, recon_regular for regular faces), so the
 * VJP is taken at the forward's own face states in either precision. */
template <int a, typename T, typename R, bool REG, bool NS>
__device__ __forceinline__ void adj_face_body(const DevMesh &M, const DevParams &P, int j,
                                              const double *__restrict__ S,
                                              const double *__restrict__ lam, double cdt,
                                              const R *__restrict__ SGm, R *__restrict__ lUL,
                                              R *__restrict__ lUR, R *__restrict__ lG) {
  const int iL = M.fL[j], iR = M.fR[j];
  const uint32_t meta = M.fmeta[j];
  const int fl = meta & 7, fr = (meta >> 3) & 7, bc = REG ? 0 : (meta >> 6) & 3,
            lev = (meta >> 8) & 0xff;
  const T s = (T)c_face_sign[fl];
  const bool first = P.ischeme == 1;
  double XL[5], XR[5];
  if (first) {
    load_S<5>(S, M.ns, iL, XL);
    if (bc == 0) load_S<5>(S, M.ns, iR, XR);
  } else if constexpr (REG) {
    recon_regular<5, T>(M, P, S, j, iL, iR, lev, a, XL, XR);
  } else {
    recon_face<5, T>(M, P, S, iL, fl, XL);
    if (bc == 0) recon_face<5, T>(M, P, S, iR, fr, XR);
  }
  /* mu = cdt lam / Vol with one quotient per side, not one per variable */
  const double cL = cdt / cell_vol(M, iL), cR = cdt / cell_vol(M, iR);
  T UL[5], UR[5], ob[5], ULb[5], URb[5];
  const T A = (T)M.area[3 * lev + a];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    UL[v] = (T)XL[v];
    UR[v] = bc == 0 ? (T)XR[v] : (bc == 1 ? (T)P.inlet[v] : UL[v]);
    const double muL = lam[(size_t)v * M.ns + iL] * cL;
    const double muR = bc == 0 ? lam[(size_t)v * M.ns + iR] * cR : 0.0;
    ob[v] = -A * (T)(muL - muR);
    ULb[v] = URb[v] = (T)0;
  }
  hllc_vjp_ideal<a, T>(P, UL, UR, s, ob, ULb, URb);
  if constexpr (NS) {
    /* a boundary face's other side is its own cell (face_R = face_L): both
     * gradient bars belong to it, which the per-cell gather counts twice */
    T gL[9], gR[9], gb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      gL[k] = (T)SGm[(size_t)k * M.ns + iL];
      gR[k] = (T)SGm[(size_t)k * M.ns + iR];
    }
    visc_vjp_ideal<a, T>(P, UL, UR, gL, gR, s, ob, ULb, URb, gb);
#pragma unroll
    for (int k = 0; k < 9; ++k) lG[(size_t)j * 9 + k] = (R)gb[k];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    if (bc == 2) { ULb[v] += URb[v]; URb[v] = (T)0; }
    if (bc == 1) URb[v] = (T)0;
    lUL[(size_t)j * 5 + v] = (R)ULb[v];
    lUR[(size_t)j * 5 + v] = (R)URb[v];
  }
}

template <typename T, typename R, bool NS>
__global__ __launch_bounds__(FV_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_face(
    DevMesh M, DevParams P, int n0, int n1, int n2, int r0, int r1, int r2,
    const double *__restrict__ S, const double *__restrict__ lam, double cdt,
    const R *__restrict__ SGm, R *__restrict__ lUL, R *__restrict__ lUR, R *__restrict__ lG) {
  /* the axis groups in order (XCD-contiguous): dispatching the irregular blocks first,
   * as the forward face kernels do, measured slower here.  The face index is formed
   * inside each branch: one index live across the branches was spilled (the kernel
   * sits at its register limit) and cost 7%. */
  const int b0 = (n0 + FV_BLOCK - 1) / FV_BLOCK, b1 = (n1 + FV_BLOCK - 1) / FV_BLOCK;
  const int lb = logical_block();
  const int t = threadIdx.x;
  if (lb < b0) {
    const int q = lb * FV_BLOCK + t;
    if (q < r0) adj_face_body<0, T, R, true, NS>(M, P, q, S, lam, cdt, SGm, lUL, lUR, lG);
    else if (q < n0) adj_face_body<0, T, R, false, NS>(M, P, q, S, lam, cdt, SGm, lUL, lUR, lG);
  } else if (lb < b0 + b1) {
    const int q = (lb - b0) * FV_BLOCK + t;
    if (q < r1) adj_face_body<1, T, R, true, NS>(M, P, n0 + q, S, lam, cdt, SGm, lUL, lUR, lG);
    else if (q < n1) adj_face_body<1, T, R, false, NS>(M, P, n0 + q, S, lam, cdt, SGm, lUL, lUR, lG);
  } else {
    const int q = (lb - b0 - b1) * FV_BLOCK + t;
    if (q < r2)
      adj_face_body<2, T, R, true, NS>(M, P, n0 + n1 + q, S, lam, cdt, SGm, lUL, lUR, lG);
    else if (q < n2)
      adj_face_body<2, T, R, false, NS>(M, P, n0 + n1 + q, S, lam, cdt, SGm, lUL, lUR, lG);
  }
}

/* bar of each cell's momentum gradients: the sum over its face slots of the
 * faces' gradient bars (twice on a boundary face, whose two sides are the cell) */
template <typename R>
__global__ __launch_bounds__(FV_BLOCK) void k_adj_sg(DevMesh M, const R *__restrict__ lG,
                                                     double *__restrict__ SGb) {
  const int i = logical_block() * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  const uint32_t cf = M.cflag[i];
  double acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int g = M.gslot[(size_t)f * M.ns + i];
    if (g < 0) continue;
    const int cls = (cf >> (2 * f)) & 3;
    const int jf = g >> 1;
    const int nk = cls == CLS_FINER ? 4 : 1;
    const double w = cls == CLS_BOUNDARY ? 2.0 : 1.0;
    for (int k = 0; k < nk; ++k)
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] += w * (double)lG[(size_t)(jf + k) * 9 + q];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) SGb[(size_t)k * M.ns + i] = acc[k];
}

/* bar of the face state SF[f] of cell i, gathered from the face(s) of slot f */
template <typename R>
__device__ __forceinline__ void sf_bar_at(const R *__restrict__ lUL, const R *__restrict__ lUR,
                                          uint32_t cf, int f, int g, double *out) {
  const int cls = (cf >> (2 * f)) & 3;
  /* the first row is loaded outside the branches (row 0 stands in for a face
   * without a slot), so that the loads of several calls are in flight together */
  const size_t j0 = g >= 0 ? (size_t)(g >> 1) : 0;
  const R *src = (cls != CLS_FINER && (g & 1)) ? lUR : lUL;
  R r0[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) r0[v] = src[j0 * 5 + v];
#pragma unroll
  for (int v = 0; v < 5; ++v) out[v] = 0.0;
  if (g < 0) return;
  if (cls == CLS_FINER) {
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] += (double)r0[v];
#pragma unroll
    for (int k = 1; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] += (double)lUL[(j0 + k) * 5 + v];
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = (double)r0[v];
  }
}

/* the same with the slot read here */
template <typename R>
__device__ __forceinline__ void sf_bar(const DevMesh &M, const R *__restrict__ lUL,
                                       const R *__restrict__ lUR, int i, uint32_t cf, int f,
                                       double *out) {
  sf_bar_at<R>(lUL, lUR, cf, f, M.gslot[(size_t)f * M.ns + i], out);
}

/* cells of a k_adj_recon workgroup: 512 consecutive cells, an 8 x 8 x 8 brick of a
 * uniform region in Morton order */
constexpr int ADJ_XB = 512;

/* Whether a same-level neighbour nb of a cell lies in the cell's own block of ADJ_XB
 * consecutive cells: the bars between such a pair go through LDS in k_adj_recon, all
 * others through the neighbour-bar arrays Pp / Pm. */
__device__ __forceinline__ bool same_block(int i, int nb) {
  return nb >= 0 && nb / ADJ_XB == i / ADJ_XB;
}

/* recon^T, part 1: own bar and the bars sent to the +/- neighbour of each axis.  A bar
 * for a same-level neighbour in the same block of ADJ_XB cells (most of them: a block
 * is a Morton brick) is handed over in LDS and added here; the others (across a block
 * boundary, across a level change) are written to Pp / Pm for k_adj_recv_stage, which
 * adds them.  `self` receives the own bar plus the bars received in the block. */
template <typename R>
__global__ __launch_bounds__(ADJ_XB) void k_adj_recon(DevMesh M, DevParams P,
                                                        const double *__restrict__ S,
                                                        const R *__restrict__ lUL,
                                                        const R *__restrict__ lUR,
                                                        double *__restrict__ self,
                                                        double *__restrict__ Pp,
                                                        double *__restrict__ Pm) {
  const int t = threadIdx.x;
  const int i = logical_block() * ADJ_XB + t;
  const bool live = i < M.N;
  const size_t pl = (size_t)5 * M.ns;
  double acc[5] = {0, 0, 0, 0, 0};
  if (P.ischeme == 1) {
    /* first order: no limiter, nothing to send */
    if (!live) return;
    const uint32_t cf = M.cflag[i];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      double b[5];
      sf_bar<R>(M, lUL, lUR, i, cf, f, b);
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += b[v];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) self[(size_t)v * M.ns + i] = acc[v];
    return;
  }
  /* the bars sent to the + and - neighbour of the current axis, by thread; two sets,
   * taken by axis parity, so that an axis's writes never meet the previous axis's reads
   * and one barrier per axis suffices (80 KB: two blocks per CU) */
  __shared__ double sPb[2][5][ADJ_XB], sMb[2][5][ADJ_XB];
  const uint32_t cf = live ? M.cflag[i] : 0u;
  double U[5] = {0, 0, 0, 0, 0};
  int n0 = 0, n1 = 0, n2 = 0, n3 = 0, n4 = 0, n5 = 0, s0 = -1, s1 = -1, s2 = -1, s3 = -1, s4 = -1,
      s5 = -1;
  if (live) {
    load_S<5>(S, M.ns, i, U);
    /* the neighbours and slots of all six faces are read before the axis loop, so that
     * each axis waits for one round of loads (its states and bars), not two */
    n0 = M.nbr[i]; n1 = M.nbr[(size_t)M.ns + i]; n2 = M.nbr[(size_t)2 * M.ns + i];
    n3 = M.nbr[(size_t)3 * M.ns + i]; n4 = M.nbr[(size_t)4 * M.ns + i];
    n5 = M.nbr[(size_t)5 * M.ns + i];
    s0 = M.gslot[i]; s1 = M.gslot[(size_t)M.ns + i]; s2 = M.gslot[(size_t)2 * M.ns + i];
    s3 = M.gslot[(size_t)3 * M.ns + i]; s4 = M.gslot[(size_t)4 * M.ns + i];
    s5 = M.gslot[(size_t)5 * M.ns + i];
  }
  /* one axis at a time (not unrolled): the three axes' 20 live values each would
   * otherwise be interleaved and cost the kernel most of its occupancy */
#pragma unroll 1
  for (int ax = 0; ax < 3; ++ax) {
    double(*sP)[ADJ_XB] = sPb[ax & 1];
    double(*sM)[ADJ_XB] = sMb[ax & 1];
    const int fp = ax == 0 ? 1 : (ax == 1 ? 2 : 5);
    const int fm = ax == 0 ? 3 : (ax == 1 ? 0 : 4);
    const int np = ax == 0 ? n1 : (ax == 1 ? n2 : n5), nm = ax == 0 ? n3 : (ax == 1 ? n0 : n4);
    const int cp = (cf >> (2 * fp)) & 3, cm = (cf >> (2 * fm)) & 3;
    const bool inP = live && cp == CLS_SAME && same_block(i, np);
    const bool inM = live && cm == CLS_SAME && same_block(i, nm);
    if (live) {
      const int gp = ax == 0 ? s1 : (ax == 1 ? s2 : s5), gm = ax == 0 ? s3 : (ax == 1 ? s0 : s4);
      double Up[5], Um[5], lp[5], lm[5];
      nb_state_at<5>(M, P, S, i, cf, fp, np, U, Up);
      nb_state_at<5>(M, P, S, i, cf, fm, nm, U, Um);
      sf_bar_at<R>(lUL, lUR, cf, fp, gp, lp);
      sf_bar_at<R>(lUL, lUR, cf, fm, gm, lm);
      double pcf, pcc, mcf, mcc, pic, mic;
      face_dist(M, cf, fp, pcf, pcc, pic);
      face_dist(M, cf, fm, mcf, mcc, mic);
      const bool bp = cp == CLS_BOUNDARY, bm = cm == CLS_BOUNDARY;
      const bool ghost_p = bp && (-1 - np) == 2;
      const bool ghost_m = bm && (-1 - nm) == 2;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double dP = Up[v] - U[v], dM = Um[v] - U[v];
        double g0, g1, h0, h1;
        limited_d_axis(P.limiter, P.va_eps, bp, bm, dP, dM, pic, mic, g0, g1, h0, h1);
        /* SF[fp] = U + pcf g(dP, dM), SF[fm] = U + mcf g(dM, dP) */
        double dPb = lp[v] * pcf * g0 + lm[v] * mcf * h1;
        double dMb = lp[v] * pcf * g1 + lm[v] * mcf * h0;
        /* a transmissive ghost equals the cell itself: its delta is identically 0 */
        if (ghost_p) dPb = 0.0;
        if (ghost_m) dMb = 0.0;
        acc[v] += lp[v] + lm[v] - dPb - dMb;
        sP[v][t] = dPb;
        sM[v][t] = dMb;
        if (!inP) Pp[ax * pl + (size_t)v * M.ns + i] = bp ? 0.0 : dPb;
        if (!inM) Pm[ax * pl + (size_t)v * M.ns + i] = bm ? 0.0 : dMb;
      }
    }
    __syncthreads();
    /* what the same-level neighbours of this block sent: my - neighbour's bar for its +
     * side, my + neighbour's for its - side */
    if (inM) {
      const int q = nm - (i - t);
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += sP[v][q];
    }
    if (inP) {
      const int q = np - (i - t);
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += sM[v][q];
    }
  }
  if (live) {
#pragma unroll
    for (int v = 0; v < 5; ++v) self[(size_t)v * M.ns + i] = acc[v];
  }
}

/* the probe objective's gradient at state S, added to lam: dt Vol dp/dS */
__device__ __forceinline__ void obj_grad_cell(const DevMesh &M, const DevParams &P,
                                              const double *__restrict__ S, int i, double dt,
                                              double *__restrict__ lam) {
  double U[5];
  load_S<5>(S, M.ns, i, U);
  const double c = dt * cell_vol(M, i) * (P.gamma - 1.0);
  const double r = U[0], u = U[1] / r, v = U[2] / r, w = U[3] / r;
  lam[i] += c * 0.5 * (u * u + v * v + w * w);
  lam[(size_t)1 * M.ns + i] -= c * u;
  lam[(size_t)2 * M.ns + i] -= c * v;
  lam[(size_t)3 * M.ns + i] -= c * w;
  lam[(size_t)4 * M.ns + i] += c;
}

/* recon^T, part 2: nu_i = own bar + bars sent to i by its face neighbours (a
 * neighbour whose delta used the mean of 4 finer cells sends 1/4 to each), then
 * the stage's reverse update:
 *   mode 0 (last stage):  lamSN  = a lam;  lam = b lam + nu
 *   mode 1:               lamSN += a lam;  lam = b lam + nu
 *   mode 2 (first stage): lam = lamSN + a lam + nu
 * objS (mode 2, optional): the step's input state S^n, whose objective term J_n the
 * reverse sweep adds next; it is added here (the same additions k_obj_grad makes, in
 * the same order), saving a launch per step. */
template <bool NS>
__global__ __launch_bounds__(FV_BLOCK, 4) void k_adj_recv_stage(DevMesh M, DevParams P,
                                                             const double *__restrict__ self,
                                                             const double *__restrict__ Pp,
                                                             const double *__restrict__ Pm,
                                                             const double *__restrict__ SGb,
                                                             double *__restrict__ lam,
                                                             double *__restrict__ lamSN, double a,
                                                             double b, int mode,
                                                             const double *__restrict__ objS,
                                                             const uint8_t *__restrict__ mask,
                                                             double dt) {
  const int i = logical_block() * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  const uint32_t cf = M.cflag[i];
  const size_t pl = (size_t)5 * M.ns;
  double acc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) acc[v] = self[(size_t)v * M.ns + i];
  /* the six neighbours and the bars a single neighbour sent through Pp / Pm are loaded
   * before any branch, so that all of them are in flight together (a same-level
   * neighbour in the cell's block handed its bar over in k_adj_recon's LDS, already in
   * `self`); the sums below keep their order */
  int nbs[6];
  double rin[6][5];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int ax = q >> 1, side = q & 1;
    const int f = side == 0 ? (ax == 0 ? 3 : (ax == 1 ? 0 : 4)) : (ax == 0 ? 1 : (ax == 1 ? 2 : 5));
    nbs[q] = M.nbr[(size_t)f * M.ns + i];
  }
  const bool muscl = P.ischeme != 1;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int ax = q >> 1, side = q & 1;
    const int f = side == 0 ? (ax == 0 ? 3 : (ax == 1 ? 0 : 4)) : (ax == 0 ? 1 : (ax == 1 ? 2 : 5));
    const int cls = (cf >> (2 * f)) & 3;
    const double *Pin = (side == 0 ? Pp : Pm) + ax * pl;
    const bool one = muscl && (cls == CLS_COARSER || (cls == CLS_SAME && !same_block(i, nbs[q])));
#pragma unroll
    for (int v = 0; v < 5; ++v) rin[q][v] = one ? Pin[(size_t)v * M.ns + nbs[q]] : 0.0;
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int fp = ax == 0 ? 1 : (ax == 1 ? 2 : 5);
    const int fm = ax == 0 ? 3 : (ax == 1 ? 0 : 4);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      /* my - side neighbours sent Pp (their + side is me); my + side ones sent Pm */
      const int f = side == 0 ? fm : fp;
      const double *Pin = (side == 0 ? Pp : Pm) + ax * pl;
      const int cls = (cf >> (2 * f)) & 3;
      const int nb = nbs[2 * ax + side];
      const double *r = rin[2 * ax + side];
      if (!muscl) {
        /* first order: no neighbour bars */
      } else if (cls == CLS_SAME) {
        if (!same_block(i, nb)) {
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v] += r[v];
        }
      } else if (cls == CLS_COARSER) {
#pragma unroll
        for (int v = 0; v < 5; ++v) acc[v] += r[v] / ((double)4);
      } else if (cls == CLS_FINER) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int q = M.quad[4 * nb + k];
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v] += Pin[(size_t)v * M.ns + q];
        }
      }
    }
  }
  if constexpr (NS) {
    /* transpose of SG_i[c][d] = sum_f sum_k (U_i + U_nb)_{1+c} n_d A / (2 Vol_i) (gradcalc.c:
     * 33-37): the cell's own term over its faces and sub-faces, plus, from each face
     * neighbour nb, the coefficient nb's gradient gives this cell (normal -n, the
     * sub-face area of the finer side, nb's volume).  A transmissive ghost is the cell
     * itself (its term counts twice); an inlet ghost is a constant. */
    const int L = (cf >> 16) & 0xff;
    const double Vi = M.vol[L];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int d = c_face_axis[f];
      const double sf = (double)c_face_sign[f];
      const int cls = (cf >> (2 * f)) & 3;
      const int nb = M.nbr[(size_t)f * M.ns + i];
      double own;
      if (cls == CLS_SAME || cls == CLS_COARSER) {
        const double A = M.area[3 * L + d];
        own = 0.5 * sf * A / Vi;
        const double cn = 0.5 * (-sf) * A / M.vol[(M.cflag[nb] >> 16) & 0xff];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[1 + c] += cn * SGb[(size_t)(3 * c + d) * M.ns + nb];
      } else if (cls == CLS_FINER) {
        const double A = M.area[3 * (L + 1) + d];
        own = 4.0 * (0.5 * sf * A / Vi);
        const double cn = 0.5 * (-sf) * A / M.vol[L + 1];
        for (int k = 0; k < 4; ++k) {
          const int q = M.quad[4 * nb + k];
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[1 + c] += cn * SGb[(size_t)(3 * c + d) * M.ns + q];
        }
      } else {
        const double A = M.area[3 * L + d];
        own = ((-1 - nb) == 1 ? 0.5 : 1.0) * sf * A / Vi;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[1 + c] += own * SGb[(size_t)(3 * c + d) * M.ns + i];
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const size_t o = (size_t)v * M.ns + i;
    const double l = lam[o];
    if (mode == 0) { lamSN[o] = a * l; lam[o] = b * l + acc[v]; }
    else if (mode == 1) { lamSN[o] += a * l; lam[o] = b * l + acc[v]; }
    else { lam[o] = lamSN[o] + a * l + acc[v]; }
  }
  if (objS && mask[i]) obj_grad_cell(M, P, objS, i, dt, lam);
}

/* double-double accumulation (Knuth's TwoSum): (ah, al) += (bh, bl).  The
 * objective is summed this way so that J(S0 + eps v) - J(S0 - eps v) is not
 * swamped by the rounding of J itself (about 1e-16 |J|, which a central
 * difference with eps = 1e-6 turns into 1e-10 |J| of noise). */
__host__ __device__ __forceinline__ void dd_add(double &ah, double &al, double bh, double bl) {
  const double s = ah + bh;
  const double bb = s - ah;
  double e = (ah - (s - bb)) + (bh - bb);
  e += al + bl;
  ah = s + e;
  al = e - (ah - s);
}

/* objective: part[4b..4b+3] = (J hi, J lo, dJ hi, dJ lo) over the probe cells of
 * block b (fixed pairwise order) */
__global__ __launch_bounds__(FV_BLOCK) void k_obj(DevMesh M, DevParams P,
                                                  const double *__restrict__ S,
                                                  const double *__restrict__ Sd,
                                                  const uint8_t *__restrict__ mask,
                                                  double *__restrict__ part) {
  __shared__ double sh[4][FV_BLOCK];
  const int i = blockIdx.x * FV_BLOCK + threadIdx.x;
  double pv = 0.0, pd = 0.0;
  if (i < M.N && mask[i]) {
    double U[5];
    load_S<5>(S, M.ns, i, U);
    const double V = cell_vol(M, i);
    const double g1 = P.gamma - 1.0;
    const double r = U[0], u = U[1] / r, v = U[2] / r, w = U[3] / r;
    const double ei = (U[4] / r) - 0.5 * (u * u + v * v + w * w);
    pv = (r * g1 * ei - P.gamma * P.pinf) * V;
    if (Sd) {
      double D[5];
      load_S<5>(Sd, M.ns, i, D);
      pd = V * g1 * (0.5 * (u * u + v * v + w * w) * D[0] - u * D[1] - v * D[2] - w * D[3] + D[4]);
    }
  }
  const int t = threadIdx.x;
  sh[0][t] = pv;
  sh[1][t] = 0.0;
  sh[2][t] = pd;
  sh[3][t] = 0.0;
  __syncthreads();
  for (int w = FV_BLOCK / 2; w > 0; w >>= 1) {
    if (t < w) {
      dd_add(sh[0][t], sh[1][t], sh[0][t + w], sh[1][t + w]);
      dd_add(sh[2][t], sh[3][t], sh[2][t + w], sh[3][t + w]);
    }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = sh[k][0];
}

/* J = sum of the block partials: one block, thread t sums partials t, t+256, ...
 * in order, then a fixed pairwise tree (the same order on every run); J and Jd
 * receive (hi, lo) */
__global__ __launch_bounds__(FV_BLOCK) void k_obj_final(const double *__restrict__ part, int nb,
                                                        double *__restrict__ J,
                                                        double *__restrict__ Jd) {
  __shared__ double sh[4][FV_BLOCK];
  const int t = threadIdx.x;
  double a = 0.0, al = 0.0, b = 0.0, bl = 0.0;
  for (int k = t; k < nb; k += FV_BLOCK) {
    dd_add(a, al, part[4 * k], part[4 * k + 1]);
    dd_add(b, bl, part[4 * k + 2], part[4 * k + 3]);
  }
  sh[0][t] = a;
  sh[1][t] = al;
  sh[2][t] = b;
  sh[3][t] = bl;
  __syncthreads();
  for (int w = FV_BLOCK / 2; w > 0; w >>= 1) {
    if (t < w) {
      dd_add(sh[0][t], sh[1][t], sh[0][t + w], sh[1][t + w]);
      dd_add(sh[2][t], sh[3][t], sh[2][t + w], sh[3][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    J[0] = sh[0][0];
    J[1] = sh[1][0];
    if (Jd) { Jd[0] = sh[2][0]; Jd[1] = sh[3][0]; }
  }
}

/* lam += dt Vol dp/dS on the probe */
__global__ void k_obj_grad(DevMesh M, DevParams P, const double *__restrict__ S,
                           const uint8_t *__restrict__ mask, double dt, double *__restrict__ lam) {
  const int i = blockIdx.x * FV_BLOCK + threadIdx.x;
  if (i >= M.N || !mask[i]) return;
  obj_grad_cell(M, P, S, i, dt, lam);
}

/* tangent of the reconstruction: SFd = Sd + dcf (g0 dPd + g1 dMd) */
__global__ __launch_bounds__(FV_BLOCK) void k_tlm_recon(DevMesh M, DevParams P,
                                                        const double *__restrict__ S,
                                                        const double *__restrict__ Sd,
                                                        double *__restrict__ SFd) {
  const int i = logical_block() * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  const uint32_t cf = M.cflag[i];
  double U[5], D[5];
  load_S<5>(S, M.ns, i, U);
  load_S<5>(Sd, M.ns, i, D);
  DevParams Pz = P;
#pragma unroll
  for (int v = 0; v < 5; ++v) Pz.inlet[v] = 0.0; /* the inlet ghost is a constant */
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int fp = ax == 0 ? 1 : (ax == 1 ? 2 : 5);
    const int fm = ax == 0 ? 3 : (ax == 1 ? 0 : 4);
    double Up[5], Um[5], Dp[5], Dm[5];
    nb_state<5>(M, P, S, i, cf, fp, U, Up);
    nb_state<5>(M, P, S, i, cf, fm, U, Um);
    nb_state<5>(M, Pz, Sd, i, cf, fp, D, Dp);
    nb_state<5>(M, Pz, Sd, i, cf, fm, D, Dm);
    double pcf, pcc, mcf, mcc, pic, mic;
    face_dist(M, cf, fp, pcf, pcc, pic);
    face_dist(M, cf, fm, mcf, mcc, mic);
    const bool bp = ((cf >> (2 * fp)) & 3) == CLS_BOUNDARY;
    const bool bm = ((cf >> (2 * fm)) & 3) == CLS_BOUNDARY;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double dP = Up[v] - U[v], dM = Um[v] - U[v];
      const double dPd = Dp[v] - D[v], dMd = Dm[v] - D[v];
      double g0, g1, h0, h1;
      limited_d_axis(P.limiter, P.va_eps, bp, bm, dP, dM, pic, mic, g0, g1, h0, h1);
      SFd[((size_t)fp * 5 + v) * M.ns + i] = D[v] + pcf * (g0 * dPd + g1 * dMd);
      SFd[((size_t)fm * 5 + v) * M.ns + i] = D[v] + mcf * (h0 * dMd + h1 * dPd);
    }
  }
}

/* tangent of the face flux by forward-mode dual numbers (NS: the momentum
 * gradients SGm and their tangents SGmd enter as dual numbers too) */
template <int a, bool NS>
__device__ __forceinline__ void tlm_face_body(const DevMesh &M, const DevParams &P, int j,
                                              const double *__restrict__ S,
                                              const double *__restrict__ Sd, const double *SF,
                                              const double *SFd, const double *__restrict__ SGm,
                                              const double *__restrict__ SGmd,
                                              double *__restrict__ Gd) {
  typedef Dual<double> D;
  const int iL = M.fL[j], iR = M.fR[j];
  const uint32_t meta = M.fmeta[j];
  const int fl = meta & 7, fr = (meta >> 3) & 7, bc = (meta >> 6) & 3, lev = (meta >> 8) & 0xff;
  const bool first = P.ischeme == 1;
  /* first order reads the cell-major S; MUSCL the face states [6][5][ns] */
  const double *X = first ? S : SF;
  const double *Xd = first ? Sd : SFd;
  const size_t oL = first ? (size_t)iL * 5 : (size_t)fl * 5 * M.ns + iL;
  const size_t oR = first ? (size_t)iR * 5 : (size_t)fr * 5 * M.ns + iR;
  const size_t st = first ? 1 : (size_t)M.ns;
  D UL[5], UR[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    UL[v] = D(X[oL + v * st], Xd[oL + v * st]);
    if (bc == 0) UR[v] = D(X[oR + v * st], Xd[oR + v * st]);
    else if (bc == 1) UR[v] = D(P.inlet[v], 0.0);
    else UR[v] = UL[v];
  }
  FState<D> L, R;
  make_state_ideal<D>(P, UL, L);
  make_state_ideal<D>(P, UR, R);
  if constexpr (NS) {
    D gL[9], gR[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      gL[k] = D(SGm[(size_t)k * M.ns + iL], SGmd[(size_t)k * M.ns + iL]);
      gR[k] = D(SGm[(size_t)k * M.ns + iR], SGmd[(size_t)k * M.ns + iR]);
    }
    vrows_axis<a, D>(gL, L.tv);
    vrows_axis<a, D>(gR, R.tv);
  }
  D out[5];
  hllc_face<1, NS, a, D>(P, L, R, D((double)c_face_sign[fl]), out);
  const double A = M.area[3 * lev + a];
#pragma unroll
  for (int v = 0; v < 5; ++v) Gd[(size_t)j * 5 + v] = -(A)*out[v].d;
}

template <bool NS>
__global__ __launch_bounds__(FV_BLOCK) void k_tlm_face(DevMesh M, DevParams P, int b0, int b1,
                                                       int n0, int n1, int n2,
                                                       const double *__restrict__ S,
                                                       const double *__restrict__ Sd,
                                                       const double *__restrict__ SF,
                                                       const double *__restrict__ SFd,
                                                       const double *__restrict__ SGm,
                                                       const double *__restrict__ SGmd,
                                                       double *__restrict__ Gd) {
  const int lb = logical_block();
  const int t = threadIdx.x;
  if (lb < b0) {
    const int j = lb * FV_BLOCK + t;
    if (j < n0) tlm_face_body<0, NS>(M, P, j, S, Sd, SF, SFd, SGm, SGmd, Gd);
  } else if (lb < b0 + b1) {
    const int q = (lb - b0) * FV_BLOCK + t;
    if (q < n1) tlm_face_body<1, NS>(M, P, n0 + q, S, Sd, SF, SFd, SGm, SGmd, Gd);
  } else {
    const int q = (lb - b0 - b1) * FV_BLOCK + t;
    if (q < n2) tlm_face_body<2, NS>(M, P, n0 + n1 + q, S, Sd, SF, SFd, SGm, SGmd, Gd);
  }
}

/* ------------------------------------------------------------------ host side */

static inline int nblk(int n) { return (n + FV_BLOCK - 1) / FV_BLOCK; }

static int adj_check(fvg_ctx *c) {
  if (!c) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  if (c->prm.model != 1)
    return set_err(c, FVG_ERR_UNSUPPORTED, "adjoint/tangent: model 1 (ideal/stiffened gas) only");
  if (c->prm.iprimtv || c->prm.wave_speeds != FVG_WAVES_DAVIS || c->prm.prandtl > 0.0)
    return set_err(c, FVG_ERR_UNSUPPORTED,
                   "adjoint/tangent: conserved-variable MUSCL, Davis speeds, no heat conduction");
  return FVG_OK;
}

static int adj_alloc(fvg_ctx *c) {
  if (c->adj) return FVG_OK;
  fv_adj_state *a = new fv_adj_state();
  c->adj = a;
  a->plane = (size_t)5 * c->ns;
  const size_t pl = a->plane * 8;
  HIPCHK(c, hipMalloc(&a->mask, (size_t)c->ns));
  HIPCHK(c, hipMemsetAsync(a->mask, 0, (size_t)c->ns, c->stream));
  HIPCHK(c, hipMalloc(&a->lam, pl));
  HIPCHK(c, hipMalloc(&a->lamSN, pl));
  HIPCHK(c, hipMalloc(&a->self, pl));
  HIPCHK(c, hipMalloc(&a->Pp, 3 * pl));
  HIPCHK(c, hipMalloc(&a->Pm, 3 * pl));
  HIPCHK(c, hipMalloc(&a->lamUL, (size_t)5 * c->nfp * c->rsize));
  HIPCHK(c, hipMalloc(&a->lamUR, (size_t)5 * c->nfp * c->rsize));
  HIPCHK(c, hipMalloc(&a->lamG, (size_t)9 * c->nfp * c->rsize));
  HIPCHK(c, hipMalloc(&a->SGb, (size_t)9 * c->ns * 8));
  HIPCHK(c, hipMemsetAsync(a->SGb, 0, (size_t)9 * c->ns * 8, c->stream));
  a->part_cap = nblk(c->N);
  HIPCHK(c, hipMalloc(&a->part, (size_t)4 * a->part_cap * 8));
  return FVG_OK;
}

/* J and (optionally) dJ contributions of the state S / tangent Sd into slot n */
static void obj_step(fvg_ctx *c, const double *S, const double *Sd, int n) {
  fv_adj_state *a = c->adj;
  DevMesh M = dev_mesh(c);
  const int nb = nblk(c->N);
  hipLaunchKernelGGL(k_obj, dim3(nb), dim3(FV_BLOCK), 0, c->stream, M, c->dp, S, Sd, a->mask,
                     a->part);
  hipLaunchKernelGGL(k_obj_final, dim3(1), dim3(FV_BLOCK), 0, c->stream, a->part, nb,
                     a->Jsteps + 2 * n, Sd ? a->Jdsteps + 2 * n : (double *)nullptr);
}

/* the reverse of one stage taken at the stage's input state S:
 * nu = R'(S)^T (cdt lam / Vol), then lam / lamSN updated per `mode`; objS: the state
 * whose objective gradient is added after the update (mode 2), or null */
static void stage_vjp(fvg_ctx *c, const double *S, double cdt, double ca, double cb, int mode,
                      const double *objS = nullptr) {
  fv_adj_state *a = c->adj;
  DevMesh M = dev_mesh(c);
  const int n0 = c->nfa[0], n1 = c->nfa[1], n2 = c->nfa[2];
  const int b0 = nblk(n0), b1 = nblk(n1), b2 = nblk(n2);
  const int r0 = c->nreg[0], r1 = c->nreg[1], r2 = c->nreg[2];
  /* grids padded to a multiple of the 8 XCDs, so that logical_block() gives each XCD a
   * contiguous range (the kernels' bounds checks drop the padding) */
  const int nc = (int)round_up(nblk(c->N), 8);
  const int nx = (int)round_up((c->N + ADJ_XB - 1) / ADJ_XB, 8);
  const bool ns = c->prm.ns != 0;
  if (ns) fv_launch_grad_raw(c, S); /* the momentum gradients of the stage's input, into c->SGm */
  dim3 gf((int)round_up(b0 + b1 + b2, 8)), b(FV_BLOCK);
  if (c->rsize == 8) {
    const double *SGm = (const double *)c->SGm;
    double *lUL = (double *)a->lamUL, *lUR = (double *)a->lamUR, *lG = (double *)a->lamG;
    if (ns)
      hipLaunchKernelGGL((k_adj_face<double, double, true>), gf, b, 0, c->stream, M, c->dp,
                         n0, n1, n2, r0, r1, r2, S, a->lam, cdt, SGm, lUL, lUR, lG);
    else
      hipLaunchKernelGGL((k_adj_face<double, double, false>), gf, b, 0, c->stream, M, c->dp,
                         n0, n1, n2, r0, r1, r2, S, a->lam, cdt, SGm, lUL, lUR, lG);
    if (ns)
      hipLaunchKernelGGL(k_adj_sg<double>, dim3(nc), b, 0, c->stream, M, (const double *)lG, a->SGb);
    hipLaunchKernelGGL((k_adj_recon<double>), dim3(nx), dim3(ADJ_XB), 0, c->stream, M, c->dp, S,
                       (const double *)lUL, (const double *)lUR, a->self, a->Pp, a->Pm);
  } else {
    const float *SGm = (const float *)c->SGm;
    float *lUL = (float *)a->lamUL, *lUR = (float *)a->lamUR, *lG = (float *)a->lamG;
    if (ns)
      hipLaunchKernelGGL((k_adj_face<float, float, true>), gf, b, 0, c->stream, M, c->dp,
                         n0, n1, n2, r0, r1, r2, S, a->lam, cdt, SGm, lUL, lUR, lG);
    else
      hipLaunchKernelGGL((k_adj_face<float, float, false>), gf, b, 0, c->stream, M, c->dp,
                         n0, n1, n2, r0, r1, r2, S, a->lam, cdt, SGm, lUL, lUR, lG);
    if (ns)
      hipLaunchKernelGGL(k_adj_sg<float>, dim3(nc), b, 0, c->stream, M, (const float *)lG, a->SGb);
    hipLaunchKernelGGL((k_adj_recon<float>), dim3(nx), dim3(ADJ_XB), 0, c->stream, M, c->dp, S,
                       (const float *)lUL, (const float *)lUR, a->self, a->Pp, a->Pm);
  }
  if (ns)
    hipLaunchKernelGGL(k_adj_recv_stage<true>, dim3(nc), b, 0, c->stream, M, c->dp, a->self, a->Pp,
                       a->Pm, (const double *)a->SGb, a->lam, a->lamSN, ca, cb, mode, objS, a->mask,
                       a->dt);
  else
    hipLaunchKernelGGL(k_adj_recv_stage<false>, dim3(nc), b, 0, c->stream, M, c->dp, a->self, a->Pp,
                       a->Pm, (const double *)a->SGb, a->lam, a->lamSN, ca, cb, mode, objS, a->mask,
                       a->dt);
}

/* Segment store: 3 ckpt + 1 states.  Step m of a segment has its stage inputs at
 * seg_state(m, 0..2) (S^n, S1, S2) and S^{n+1} at seg_state(m + 1, 0); the first
 * state of a segment is its checkpoint, read in place. */
static double *seg_state(fv_adj_state *a, int sg, int m, int k) {
  if (m == 0 && k == 0) return a->ckpts + (size_t)sg * a->plane;
  return a->seg + (size_t)(3 * m + k) * a->plane;
}

/* the SSP-RK3 steps n0 .. n1-1 of segment sg from its checkpoint into the store.
 * The reverse sweep of a segment reads the stage inputs S^n, S1, S2 of its steps; the
 * state after its last step is the next checkpoint, so a recomputation (`end` false)
 * skips that step's third stage, except in the last segment, whose final state carries
 * the objective's last term. */
static void seg_forward(fvg_ctx *c, int sg, int n0, int n1, bool objective, bool end) {
  fv_adj_state *a = c->adj;
  for (int n = n0; n < n1; ++n) {
    const int m = n - n0;
    double *X0 = seg_state(a, sg, m, 0);
    fv_stage_ptrs(c, X0, X0, seg_state(a, sg, m, 1), 1, a->dt);
    fv_stage_ptrs(c, seg_state(a, sg, m, 1), X0, seg_state(a, sg, m, 2), 2, a->dt);
    if (n == n1 - 1 && !end) break;
    fv_stage_ptrs(c, seg_state(a, sg, m, 2), X0, seg_state(a, sg, m + 1, 0), 3, a->dt);
    if (objective) obj_step(c, seg_state(a, sg, m + 1, 0), nullptr, n);
  }
}

/* sum_n dt (hi_n + lo_n) in double-double: dt hi_n exactly as a product and its
 * fma rounding error */
static void dt_sum(double dt, const std::vector<double> &hl, int nsteps, double *hi, double *lo) {
  double sh = 0.0, sl = 0.0;
  for (int n = 0; n < nsteps; ++n) {
    const double p = dt * hl[2 * n];
    const double e = std::fma(dt, hl[2 * n], -p) + dt * hl[2 * n + 1];
    dd_add(sh, sl, p, e);
  }
  *hi = sh;
  *lo = sl;
}

extern "C" {

int fvg_adj_set_probe(fvg_ctx *c, const uint8_t *mask) {
  int rc = adj_check(c);
  if (rc) return rc;
  if (!mask) return set_err(c, FVG_ERR_ARG, "null mask");
  rc = adj_alloc(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->adj->mask, mask, (size_t)c->N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

int fvg_adj_forward(fvg_ctx *c, int nsteps, int ckpt, double dt, double *J) {
  int rc = adj_check(c);
  if (rc) return rc;
  if (nsteps < 1 || ckpt < 0 || !J) return set_err(c, FVG_ERR_ARG, "bad nsteps/ckpt");
  (void)hipSetDevice(c->device);
  rc = adj_alloc(c);
  if (rc) return rc;
  fv_adj_state *a = c->adj;
  const size_t pl = a->plane * 8;
  if (a->ckpts) (void)hipFree(a->ckpts);
  if (a->seg) (void)hipFree(a->seg);
  if (a->Jsteps) (void)hipFree(a->Jsteps);
  a->ckpts = a->seg = a->Jsteps = nullptr;
  a->fwd_done = false;
  a->seg_valid = -1;
  if (ckpt == 0) {
    /* the widest spacing whose stores, (3 ckpt + 1) segment states and one checkpoint per
     * segment, fit in half the free device memory: the whole rollout where it fits, so
     * that the reverse sweep recomputes nothing */
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    const size_t budget = fr / 2;
    ckpt = 1;
    for (int k = nsteps; k >= 1; --k) {
      const size_t need = ((size_t)3 * k + 1 + (size_t)(nsteps + k - 1) / k) * pl;
      if (need <= budget) { ckpt = k; break; }
    }
  }
  if (ckpt > nsteps) ckpt = nsteps;
  const int nck = (nsteps + ckpt - 1) / ckpt;
  HIPCHK(c, hipMalloc(&a->ckpts, (size_t)nck * pl));
  HIPCHK(c, hipMalloc(&a->seg, (size_t)(3 * ckpt + 1) * pl));
  HIPCHK(c, hipMalloc(&a->Jsteps, (size_t)2 * nsteps * 8));
  a->nsteps = nsteps;
  a->ckpt = ckpt;
  a->nck = nck;
  a->dt = dt;
  /* every segment but the last runs in place on S, storing only its checkpoint;
   * the last one runs in the segment store, which the backward sweep then reads
   * without recomputing it */
  const int nlast = (nck - 1) * ckpt;
  for (int n = 0; n < nlast; ++n) {
    if (n % ckpt == 0)
      HIPCHK(c, hipMemcpyAsync(a->ckpts + (size_t)(n / ckpt) * a->plane, c->S, pl,
                               hipMemcpyDeviceToDevice, c->stream));
    for (int k = 1; k <= 3; ++k) fv_stage(c, k, dt, k == 1, false);
    obj_step(c, c->S, nullptr, n);
  }
  HIPCHK(c, hipMemcpyAsync(a->ckpts + (size_t)(nck - 1) * a->plane, c->S, pl,
                           hipMemcpyDeviceToDevice, c->stream));
  seg_forward(c, nck - 1, nlast, nsteps, true, true);
  a->seg_valid = nck - 1;
  HIPCHK(c, hipMemcpyAsync(c->S, seg_state(a, nck - 1, nsteps - nlast, 0), pl,
                           hipMemcpyDeviceToDevice, c->stream));
  c->sf_valid = false;
  c->stage_in = nullptr;
  std::vector<double> js(2 * nsteps);
  HIPCHK(c, hipMemcpyAsync(js.data(), a->Jsteps, js.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dt_sum(dt, js, nsteps, &a->J_hi, &a->J_lo);
  *J = a->J_hi;
  a->fwd_done = true;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  return FVG_OK;
}

int fvg_adj_checkpoint_spacing(const fvg_ctx *c) {
  if (!c) return FVG_ERR_ARG;
  return (c->adj && c->adj->fwd_done) ? c->adj->ckpt : 0;
}

int fvg_adj_backward(fvg_ctx *c, double *grad_host, void *grad_dev) {
  int rc = adj_check(c);
  if (rc) return rc;
  fv_adj_state *a = c->adj;
  if (!a || !a->fwd_done) return set_err(c, FVG_ERR_STATE, "fvg_adj_forward first");
  (void)hipSetDevice(c->device);
  DevMesh M = dev_mesh(c);
  const size_t pl = a->plane * 8;
  const int nc = nblk(c->N);
  const double dt = a->dt;
  /* S_k = a SN + b S_{k-1} + c dt/Vol R(S_{k-1}) for stages k = 1, 2, 3 */
  const double cst[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
  const double ast[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0};
  const double bst[3] = {0.0, 1.0 / 4.0, 2.0 / 3.0};
  HIPCHK(c, hipMemsetAsync(a->lam, 0, pl, c->stream));
  for (int sg = a->nck - 1; sg >= 0; --sg) {
    const int n0 = sg * a->ckpt, n1 = std::min(n0 + a->ckpt, a->nsteps);
    if (a->seg_valid != sg) {
      seg_forward(c, sg, n0, n1, false, sg == a->nck - 1);
      a->seg_valid = sg;
    }
    for (int n = n1 - 1; n >= n0; --n) {
      const int m = n - n0;
      /* d J_{n+1} / d S^{n+1}: for the last step here; for the others it was added by
       * the first-stage reverse of step n + 1, whose input state is S^{n+1} */
      if (n == a->nsteps - 1)
        hipLaunchKernelGGL(k_obj_grad, dim3(nc), dim3(FV_BLOCK), 0, c->stream, M, c->dp,
                           (const double *)seg_state(a, sg, m + 1, 0), a->mask, dt, a->lam);
      for (int k = 3; k >= 1; --k)
        stage_vjp(c, seg_state(a, sg, m, k - 1), cst[k - 1] * dt, ast[k - 1], bst[k - 1],
                  k == 3 ? 0 : (k == 2 ? 1 : 2),
                  (k == 1 && n >= 1) ? (const double *)seg_state(a, sg, m, 0) : nullptr);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  for (int v = 0; v < 5; ++v) {
    if (grad_host)
      HIPCHK(c, hipMemcpyAsync(grad_host + (size_t)v * c->N, a->lam + (size_t)v * c->ns,
                               (size_t)c->N * 8, hipMemcpyDeviceToHost, c->stream));
    if (grad_dev)
      HIPCHK(c, hipMemcpyAsync((double *)grad_dev + (size_t)v * c->N, a->lam + (size_t)v * c->ns,
                               (size_t)c->N * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

int fvg_tlm(fvg_ctx *c, const double *dir, int nsteps, double dt, double *J, double *dJ) {
  int rc = adj_check(c);
  if (rc) return rc;
  if (!dir || nsteps < 1 || !J || !dJ) return set_err(c, FVG_ERR_ARG, "bad tangent arguments");
  if (c->rsize != 8) return set_err(c, FVG_ERR_UNSUPPORTED, "tangent runs in plain FP64");
  (void)hipSetDevice(c->device);
  rc = adj_alloc(c);
  if (rc) return rc;
  fv_adj_state *a = c->adj;
  const size_t pl = a->plane * 8;
  if (!a->Sd) {
    HIPCHK(c, hipMalloc(&a->Sd, pl));
    HIPCHK(c, hipMalloc(&a->SNd, pl));
    HIPCHK(c, hipMalloc(&a->SFd, 6 * pl));
    HIPCHK(c, hipMalloc(&a->Gd, (size_t)5 * c->nfp * 8));
    HIPCHK(c, hipMalloc(&a->SGmd, (size_t)9 * c->ns * 8));
    HIPCHK(c, hipMemsetAsync(a->SFd, 0, 6 * pl, c->stream));
  }
  if (a->Jsteps) (void)hipFree(a->Jsteps);
  if (a->Jdsteps) (void)hipFree(a->Jdsteps);
  a->Jsteps = a->Jdsteps = nullptr;
  HIPCHK(c, hipMalloc(&a->Jsteps, (size_t)2 * nsteps * 8));
  HIPCHK(c, hipMalloc(&a->Jdsteps, (size_t)2 * nsteps * 8));
  HIPCHK(c, hipMemsetAsync(a->Sd, 0, pl, c->stream));
  {
    /* the tangent state is cell-major like S */
    std::vector<double> t((size_t)5 * c->N);
    for (int v = 0; v < 5; ++v)
      for (int i = 0; i < c->N; ++i) t[(size_t)i * 5 + v] = dir[(size_t)v * c->N + i];
    HIPCHK(c, hipMemcpyAsync(a->Sd, t.data(), t.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  DevMesh M = dev_mesh(c);
  const int nc = nblk(c->N);
  const int n0 = c->nfa[0], n1 = c->nfa[1], n2 = c->nfa[2];
  const int b0 = nblk(n0), b1 = nblk(n1), b2 = nblk(n2);
  for (int n = 0; n < nsteps; ++n) {
    for (int k = 1; k <= 3; ++k) {
      if (c->prm.ischeme == 2) {
        fv_launch_recon(c, c->S);
        hipLaunchKernelGGL(k_tlm_recon, dim3(nc), dim3(FV_BLOCK), 0, c->stream, M, c->dp, c->S, a->Sd,
                           a->SFd);
      }
      if (c->prm.ns) {
        fv_launch_grad_raw(c, c->S);
        fv_launch_grad_tangent(c, a->Sd, a->SGmd);
        hipLaunchKernelGGL(k_tlm_face<true>, dim3(b0 + b1 + b2), dim3(FV_BLOCK), 0, c->stream, M,
                           c->dp, b0, b1, n0, n1, n2, c->S, a->Sd, (const double *)c->SF, a->SFd,
                           (const double *)c->SGm, (const double *)a->SGmd, a->Gd);
      } else {
        hipLaunchKernelGGL(k_tlm_face<false>, dim3(b0 + b1 + b2), dim3(FV_BLOCK), 0, c->stream, M,
                           c->dp, b0, b1, n0, n1, n2, c->S, a->Sd, (const double *)c->SF, a->SFd,
                           (const double *)nullptr, (const double *)nullptr, a->Gd);
      }
      /* tangent stage update is the same linear combination (k_gather_rk on tangents) */
      fv_launch_gather_ptrs(c, a->Gd, a->Sd, a->SNd, k, dt, k == 1);
      fv_stage(c, k, dt, k == 1, false);
    }
    obj_step(c, c->S, a->Sd, n);
  }
  std::vector<double> js(2 * nsteps), jd(2 * nsteps);
  HIPCHK(c, hipMemcpyAsync(js.data(), a->Jsteps, js.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(jd.data(), a->Jdsteps, jd.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  double jl, dl;
  dt_sum(dt, js, nsteps, J, &jl);
  dt_sum(dt, jd, nsteps, dJ, &dl);
  a->J_hi = *J;
  a->J_lo = jl;
  return FVG_OK;
}

int fvg_adj_objective(const fvg_ctx *c, double *J_hi_lo) {
  if (!c || !J_hi_lo) return FVG_ERR_ARG;
  if (!c->adj || !c->adj->Jsteps) return FVG_ERR_STATE;
  J_hi_lo[0] = c->adj->J_hi;
  J_hi_lo[1] = c->adj->J_lo;
  return FVG_OK;
}

/* host-side derivative checks of the face flux (no GPU needed) */
int fvg_face_flux_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                       double *out) {
  if (!p || axis < 0 || axis > 2 || !UL || !UR || !out || p->model != 1) return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  P.gamma = p->gamma;
  P.pinf = p->pinf;
  FState<double> L, R;
  make_state_ideal<double>(P, UL, L);
  make_state_ideal<double>(P, UR, R);
  const double s = sign < 0 ? -1.0 : 1.0;
  if (axis == 0) hllc_face<1, false, 0, double>(P, L, R, s, out);
  else if (axis == 1) hllc_face<1, false, 1, double>(P, L, R, s, out);
  else hllc_face<1, false, 2, double>(P, L, R, s, out);
  return FVG_OK;
}

int fvg_face_vjp_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                      const double *ob, double *ULb, double *URb) {
  if (!p || axis < 0 || axis > 2 || !UL || !UR || !ob || !ULb || !URb || p->model != 1)
    return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  P.gamma = p->gamma;
  P.pinf = p->pinf;
  for (int v = 0; v < 5; ++v) ULb[v] = URb[v] = 0.0;
  const double s = sign < 0 ? -1.0 : 1.0;
  if (axis == 0) hllc_vjp_ideal<0, double>(P, UL, UR, s, ob, ULb, URb);
  else if (axis == 1) hllc_vjp_ideal<1, double>(P, UL, UR, s, ob, ULb, URb);
  else hllc_vjp_ideal<2, double>(P, UL, UR, s, ob, ULb, URb);
  return FVG_OK;
}

int fvg_face_jvp_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                      const double *ULd, const double *URd, double *outd) {
  if (!p || axis < 0 || axis > 2 || !UL || !UR || !ULd || !URd || !outd || p->model != 1)
    return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  P.gamma = p->gamma;
  P.pinf = p->pinf;
  typedef Dual<double> D;
  D uL[5], uR[5], o[5];
  for (int v = 0; v < 5; ++v) { uL[v] = D(UL[v], ULd[v]); uR[v] = D(UR[v], URd[v]); }
  FState<D> L, R;
  make_state_ideal<D>(P, uL, L);
  make_state_ideal<D>(P, uR, R);
  const D s(sign < 0 ? -1.0 : 1.0);
  if (axis == 0) hllc_face<1, false, 0, D>(P, L, R, s, o);
  else if (axis == 1) hllc_face<1, false, 1, D>(P, L, R, s, o);
  else hllc_face<1, false, 2, D>(P, L, R, s, o);
  for (int v = 0; v < 5; ++v) outd[v] = o[v].d;
  return FVG_OK;
}

/* the Navier-Stokes face flux (HLLC + viscous flux; gL, gR the 9 momentum gradients
 * of each side) and its VJP, including the bars of the gradients (gb, per side) */
int fvg_face_flux_ns_host(const fvg_params *p, int axis, int sign, const double *UL,
                          const double *UR, const double *gL, const double *gR, double *out) {
  if (!p || axis < 0 || axis > 2 || !UL || !UR || !gL || !gR || !out || p->model != 1)
    return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  P.gamma = p->gamma;
  P.pinf = p->pinf;
  P.mu = p->mu;
  P.visc_work = p->visc_work;
  FState<double> L, R;
  make_state_ideal<double>(P, UL, L);
  make_state_ideal<double>(P, UR, R);
  const double s = sign < 0 ? -1.0 : 1.0;
  if (axis == 0) {
    vrows_axis<0, double>(gL, L.tv);
    vrows_axis<0, double>(gR, R.tv);
    hllc_face<1, true, 0, double>(P, L, R, s, out);
  } else if (axis == 1) {
    vrows_axis<1, double>(gL, L.tv);
    vrows_axis<1, double>(gR, R.tv);
    hllc_face<1, true, 1, double>(P, L, R, s, out);
  } else {
    vrows_axis<2, double>(gL, L.tv);
    vrows_axis<2, double>(gR, R.tv);
    hllc_face<1, true, 2, double>(P, L, R, s, out);
  }
  return FVG_OK;
}

int fvg_face_vjp_ns_host(const fvg_params *p, int axis, int sign, const double *UL,
                         const double *UR, const double *gL, const double *gR, const double *ob,
                         double *ULb, double *URb, double *gb) {
  if (!p || axis < 0 || axis > 2 || !UL || !UR || !gL || !gR || !ob || !ULb || !URb || !gb ||
      p->model != 1)
    return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  P.gamma = p->gamma;
  P.pinf = p->pinf;
  P.mu = p->mu;
  P.visc_work = p->visc_work;
  for (int v = 0; v < 5; ++v) ULb[v] = URb[v] = 0.0;
  const double s = sign < 0 ? -1.0 : 1.0;
  if (axis == 0) {
    hllc_vjp_ideal<0, double>(P, UL, UR, s, ob, ULb, URb);
    visc_vjp_ideal<0, double>(P, UL, UR, gL, gR, s, ob, ULb, URb, gb);
  } else if (axis == 1) {
    hllc_vjp_ideal<1, double>(P, UL, UR, s, ob, ULb, URb);
    visc_vjp_ideal<1, double>(P, UL, UR, gL, gR, s, ob, ULb, URb, gb);
  } else {
    hllc_vjp_ideal<2, double>(P, UL, UR, s, ob, ULb, URb);
    visc_vjp_ideal<2, double>(P, UL, UR, gL, gR, s, ob, ULb, URb, gb);
  }
  return FVG_OK;
}

int fvg_limiter_host(int limiter, double eps, double d0, double d1, double dp1cc, double dm1cc,
                     double *g, double *g0, double *g1) {
  if (!g || !g0 || !g1) return FVG_ERR_ARG;
  *g = limited_d(limiter, eps, false, d0, d1, dp1cc, dm1cc, *g0, *g1);
  return FVG_OK;
}

int fvg_limiter_axis_host(int limiter, double eps, int bp, int bm, double dP, double dM,
                          double ip, double im, double *out) {
  if (!out) return FVG_ERR_ARG;
  limited_d_axis(limiter, eps, bp != 0, bm != 0, dP, dM, ip, im, out[0], out[1], out[2], out[3]);
  return FVG_OK;
}

}  // extern "C"
