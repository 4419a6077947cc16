/*
 * Private declarations shared by the HIP translation units: the device mesh
 * view, the context, and the cell helpers of the reconstruction (neighbour
 * states, distance() factors, limiter).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <utility>
#include <vector>

#include "fvgpu.h"
#include "fv_physics.h"

/* ------------------------------------------------------------------ constants */

static __constant__ int c_face_axis[6] = {1, 0, 1, 0, 2, 2};
static __constant__ int c_face_sign[6] = {-1, 1, 1, -1, -1, 1};


#define FV_BLOCK 256
#define FV_NKT 8

enum { CLS_SAME = 0, CLS_FINER = 1, CLS_COARSER = 2, CLS_BOUNDARY = 3 };

static inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
/* host copy of c_face_axis */
static inline int c_face_axis_host(int f) { return f == 1 || f == 3 ? 0 : (f == 0 || f == 2 ? 1 : 2); }

/* ------------------------------------------------------------------ mesh views */

struct DevMesh {
  int N, ns, F, nfp, Q, nlevel;
  const uint32_t *cflag;
  const int32_t *nbr;      /* [6][ns] */
  const int32_t *quad;     /* [Q][4] */
  const int32_t *fL, *fR;
  const int32_t *fLL, *fRR; /* far-side neighbours of regular faces */
  const uint32_t *fmeta;
  const int32_t *gslot;    /* [6][ns] */
  const double *h, *dist, *area, *vol, *idist;
};

/* ------------------------------------------------------------------ helpers */

/* The conserved-state arrays S, SN, RHS are cell-major (array of structs):
 * S[i*NEQ + v], so the NEQ values of a gathered neighbour share one or two cache
 * lines.  The face flux G is face-major: G[j*NEQ + v]. */
template <int NEQ>
__device__ __forceinline__ void load_S(const double *__restrict__ S, int ns, int i, double *U) {
  (void)ns;
  const double *p = S + (size_t)i * NEQ;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) U[v] = p[v];
}

/* the same from the FP32 copy of the state (mixed form) */
template <int NEQ>
__device__ __forceinline__ void load_S(const float *__restrict__ S, int ns, int i, double *U) {
  (void)ns;
  const float *p = S + (size_t)i * NEQ;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) U[v] = (double)p[v];
}

/* neighbour state across face f for MUSCL deltas (facevalues_MUSCL.c:18-49 and
 * the bc ghost of deltavalues.c:12-13) */
template <int NEQ>
__device__ __forceinline__ void nb_state_at(const DevMesh &M, const DevParams &P,
                                            const double *__restrict__ S, int i, uint32_t cf,
                                            int f, int nb, const double *Uown, double *Un) {
  int cls = (cf >> (2 * f)) & 3;
  /* the single-neighbour row is loaded outside the class branches (the cell's own
   * row stands in for the other classes), so that the loads of several calls are in
   * flight together */
  double Ur[NEQ];
  load_S<NEQ>(S, M.ns, (cls == CLS_SAME || cls == CLS_COARSER) ? nb : i, Ur);
  if (cls == CLS_SAME || cls == CLS_COARSER) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) Un[v] = Ur[v];
  } else if (cls == CLS_FINER) {
    int q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = M.quad[4 * nb + k];
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc += S[(size_t)q[k] * NEQ + v] / ((double)4);
      Un[v] = acc;
    }
  } else {
    int bc = -1 - nb;
    if (bc == 1) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Un[v] = P.inlet[v];
    } else {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Un[v] = Uown[v];
    }
  }
}

/* the same with the neighbour index read here */
template <int NEQ>
__device__ __forceinline__ void nb_state(const DevMesh &M, const DevParams &P,
                                         const double *__restrict__ S, int i, uint32_t cf, int f,
                                         const double *Uown, double *Un) {
  nb_state_at<NEQ>(M, P, S, i, cf, f, M.nbr[(size_t)f * M.ns + i], Uown, Un);
}

__device__ __forceinline__ void face_dist(const DevMesh &M, uint32_t cf, int f, double &dcf,
                                          double &dcc, double &idcc) {
  int cls = (cf >> (2 * f)) & 3;
  if (cls == CLS_BOUNDARY) { dcf = 0.0; dcc = 0.0; idcc = __builtin_inf(); return; }
  int L = (cf >> 16) & 0xff;
  int a = c_face_axis[f];
  int c = cls == CLS_SAME ? 0 : (cls == CLS_FINER ? 1 : 2);
  const int k = (L * 3 + a) * 3 + c;
  dcf = M.dist[2 * k];
  dcc = M.dist[2 * k + 1];
  idcc = M.idist[k];
}

/* reconctruct_V3.c:11-45 for one variable, with the distances d_cc of the two
 * faces given as reciprocals (ip = 1/d_cc(f), im = 1/d_cc(fo)); a boundary's
 * 1/0 = inf keeps the reference's 0/0 -> NaN -> no slope and inf -> the finite
 * slope.  lim 1 = smooth van Albada (extension). */
template <typename T>
__device__ __forceinline__ T limited(int lim, double eps, bool bnd, T d0, T d1, T ip, T im) {
  T grad = (T)0;
  if (bnd) return grad;
  if (lim == FVG_MINMOD) {
    if constexpr (sizeof(T) == 4) {
      /* FP32: both tests as selects (no divergent branch), a and b formed either way
       * and only picked when |d0| > 1e-10 and a b > 0 (false for NaN); measured
       * faster than the branches in FP32, slower in FP64.  Same values either way. */
      const T a = -d1 * im, b = d0 * ip;
      const T prod = a * b;
      const T pick = fabs(a) > fabs(b) ? b : a;
      grad = (fabs(d0) > (T)1e-10 && prod > (T)0) ? pick : grad;
    } else {
      if (fabs(d0) > (T)1e-10) {
        T a = -d1 * im, b = d0 * ip;
        T prod = a * b;
        if (prod > (T)0) grad = fabs(a) > fabs(b) ? b : a;
      }
    }
  } else {
    const T e = (T)eps;
    T a = isinf(im) ? (T)0 : -d1 * im;
    T b = isinf(ip) ? (T)0 : d0 * ip;
    /* (a (b^2 + e) + b (a^2 + e)) / (a^2 + b^2 + 2 e) with every fused multiply-add
     * written out: left to the compiler, which of the two products of the numerator
     * is fused depends on the surrounding code, and the face-list and LDS-tiled
     * kernels, which must agree bit for bit, inline this in different code */
    const T num = fma(a, fma(b, b, e), b * fma(a, a, e));
    const T den = fma(a, a, fma(b, b, (T)2 * e));
    grad = qdiv(num, den);
  }
  return grad;
}

/* limited<T> on a regular face: both one-sided slopes over the same finite distance,
 * given as its reciprocal ic.  In FP32 the minmod selection is one median: med3(a, b, 0)
 * is the slope of smaller magnitude when a = -d1 ic and b = d0 ic have one sign and 0
 * when their signs differ or one is 0, the element the select form of limited<float>
 * picks (the two differ only where a b underflows, |d1| below ~1e-37).  FP64 and van
 * Albada are limited<T> itself. */
template <typename T>
__device__ __forceinline__ T limited_regular(int lim, double eps, T d0, T d1, T ic) {
  if constexpr (sizeof(T) == 4) {
    if (lim == FVG_MINMOD) {
      const float m = __builtin_amdgcn_fmed3f(-d1 * ic, d0 * ic, 0.0f);
      return fabsf(d0) > 1e-10f ? m : 0.0f;
    }
  }
  return limited<T>(lim, eps, false, d0, d1, ic, ic);
}

/* An FP32 MUSCL face value on a regular face: own + fp g with g = limited_regular(other -
 * own, far - own).  For minmod the distance's reciprocal is taken out of the median
 * (med3(-d1 ic, d0 ic, 0) = ic med3(-d1, d0, 0), ic > 0) and folded into fp: fpic = fp ic,
 * one product per face instead of two per variable.  The mixed ideal-gas kernels, the
 * LDS-tiled one and the face-list one, both form their regular face states with it. */
__device__ __forceinline__ float muscl_regular_f(int lim, double eps, float own, float other,
                                                 float far, float fp, float ic, float fpic) {
  const float d0 = other - own, d1 = far - own;
  if (lim == FVG_MINMOD) {
    const float m = __builtin_amdgcn_fmed3f(-d1, d0, 0.0f);
    return fmaf(fpic, fabsf(d0) > 1e-10f ? m : 0.0f, own);
  }
  return own + fp * limited<float>(lim, eps, false, d0, d1, ic, ic);
}

/* The update of SSP-RK stage irk (RK.c:24-47): S_k = a SN + b S + c dt R / Vol.  Every
 * kernel that applies it (the fused gather, the tiled stage, k_rk) calls this one
 * sequence of roundings with no fused multiply-adds, so that they give the same bits
 * whatever code surrounds them (the host restatement rounds the same way). */
__device__ __forceinline__ double rk_update(int irk, double s, double sn, double r, double dt,
                                            double Vol) {
#pragma clang fp contract(off)
  switch (irk) {
    case 1: return sn + dt * (1.0 / Vol) * r;
    case 2: return (3.0 / 4.0) * sn + (1.0 / 4.0) * s + (1.0 / 4.0) * dt * r / Vol;
    case 3: return (1.0 / 3.0) * sn + (2.0 / 3.0) * s + (2.0 / 3.0) * dt * r / Vol;
    case 4: return (1.0 / 2.0) * sn + (1.0 / 2.0) * s + (1.0 / 2.0) * dt * r / Vol;
    default: return s;
  }
}

/* XCD-aware block order: consecutive logical blocks (contiguous Morton ranges
 * of cells or faces) land on the same XCD and so share its L2. */
__device__ __forceinline__ int logical_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = nb >> 3;
  if (per == 0 || (nb & 7)) return b;
  return (b & 7) * per + (b >> 3);
}

/* Face blocks: [0, b0) take the x group, [b0, b0 + b1) the y group, the rest the z
 * group (the grid is b0 + b1 + b2 blocks).  With the XCD-aware block order each XCD
 * works through a stretch of one group.  (Interleaving the three groups block by block,
 * so that one XCD reads a cell range for all three axes, fetches fewer bytes but ran
 * slower on 2M cells.)  Returns the axis, blk its block in that group. */
__device__ __forceinline__ int axis_chunk(int lb, int b0, int b1, int &blk) {
  if (lb < b0) { blk = lb; return 0; }
  if (lb < b0 + b1) { blk = lb - b0; return 1; }
  blk = lb - b0 - b1;
  return 2;
}

/* Face blocks of a launch over the three axis groups, group a holding n_a faces of
 * which the first r_a are regular (ceil(n_a / FV_BLOCK) blocks per group, the grid
 * their sum).  The blocks holding any irregular face (domain boundary, hanging face,
 * or a stencil across a level change: the longest chains of dependent loads) are
 * dispatched first, so that they start early and land round-robin on all XCDs; the
 * purely regular blocks follow in the XCD-contiguous order of logical_block.  (In
 * group order the irregular tails of the three groups fell on three of the eight XCDs
 * and started last.)  Returns the axis; blk is the block's index within its group. */
__device__ __forceinline__ int face_block(int n0, int n1, int n2, int r0, int r1, int r2, int &blk) {
  const int b = blockIdx.x;
  const int rb0 = r0 / FV_BLOCK, rb1 = r1 / FV_BLOCK, rb2 = r2 / FV_BLOCK;
  const int ib0 = (n0 + FV_BLOCK - 1) / FV_BLOCK - rb0;
  const int ib1 = (n1 + FV_BLOCK - 1) / FV_BLOCK - rb1;
  const int ib2 = (n2 + FV_BLOCK - 1) / FV_BLOCK - rb2;
  const int nib = ib0 + ib1 + ib2;
  if (b < nib) {
    if (b < ib0) { blk = rb0 + b; return 0; }
    if (b < ib0 + ib1) { blk = rb1 + (b - ib0); return 1; }
    blk = rb2 + (b - ib0 - ib1);
    return 2;
  }
  const int nrb = rb0 + rb1 + rb2, c = b - nib, per = nrb >> 3;
  const int lb = (per == 0 || (nrb & 7)) ? c : (c & 7) * per + (c >> 3);
  if (lb < rb0) { blk = lb; return 0; }
  if (lb < rb0 + rb1) { blk = lb - rb0; return 1; }
  blk = lb - rb0 - rb1;
  return 2;
}

/* theta = p / rho of an ideal / stiffened-gas state, (gamma - 1) e_int - gamma p_inf / rho
 * (the p of facevalues.c:40 divided by rho), without the sound speed.  The fused
 * multiply-adds are written out so that the gradient kernel and the LDS-tiled one,
 * which inline this in different code, agree bit for bit. */
__device__ __forceinline__ double theta_ideal(const DevParams &P, const double *U) {
  const double ir = 1.0 / U[0];
  const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const double q = fma(u, u, fma(v, v, w * w));
  const double ei = fma(-0.5, q, U[4] * ir);
  return fma(P.gamma - 1.0, ei, -(P.gamma * P.pinf * ir));
}

/* whether face block blk of axis group ax is one the LDS-tiled stage covers (skip: one
 * flag per block, the groups' blocks in axis order; null: none is) */
__device__ __forceinline__ bool skipped_face_block(const uint8_t *skip, int n0, int n1, int ax,
                                                   int blk) {
  if (!skip) return false;
  const int b0 = (n0 + FV_BLOCK - 1) / FV_BLOCK, b1 = (n1 + FV_BLOCK - 1) / FV_BLOCK;
  return skip[(ax == 0 ? 0 : (ax == 1 ? b0 : b0 + b1)) + blk] != 0;
}

/* ------------------------------------------------------------------ context */

/* MUSCL face state of cell c on face f (faceinterp.c:33-44 for one face).  The
 * limited slope is computed in T (FP32 in the mixed form, from FP64 deltas) and the
 * face state is the FP64 cell state plus that increment, so the FP64 equation of
 * state sees an unrounded density (the barotropic EOS turns a density error into
 * a pressure error c^2 times larger). */
template <int NEQ, typename T>
__device__ __forceinline__ void recon_face(const DevMesh &M, const DevParams &P,
                                           const double *__restrict__ S, int c, int f,
                                           double *out) {
  const uint32_t cf = M.cflag[c];
  const int fo = f ^ (f < 4 ? 2 : 1); /* fcopp: 0<->2, 1<->3, 4<->5 */
  double U[NEQ], U0[NEQ], U1[NEQ];
  load_S<NEQ>(S, M.ns, c, U);
  nb_state<NEQ>(M, P, S, c, cf, f, U, U0);
  nb_state<NEQ>(M, P, S, c, cf, fo, U, U1);
  double pcf, pcc, mcf, mcc, pic, mic;
  face_dist(M, cf, f, pcf, pcc, pic);
  face_dist(M, cf, fo, mcf, mcc, mic);
  const bool bnd = ((cf >> (2 * f)) & 3) == CLS_BOUNDARY;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const T g = limited<T>(P.limiter, P.va_eps, bnd, (T)(U0[v] - U[v]), (T)(U1[v] - U[v]), (T)pic,
                           (T)mic);
    out[v] = U[v] + (double)((T)pcf * g);
  }
}

/* MUSCL face states of a regular face j (interior; stencil M.fLL[j], iL, iR,
 * M.fRR[j] of one level along axis a): recon_face for both sides with the
 * neighbours and the same-level distances known, the same arithmetic without
 * the class and neighbour lookups */
template <int NEQ, typename T>
__device__ __forceinline__ void recon_regular(const DevMesh &M, const DevParams &P,
                                              const double *__restrict__ S, int j, int iL, int iR,
                                              int lev, int a, double *UL, double *UR) {
  const int k = (lev * 3 + a) * 3; /* same-level distances of this level and axis */
  const double pcf = M.dist[2 * k], ic = M.idist[k];
  double A0[NEQ], A1[NEQ], B0[NEQ], B1[NEQ];
  load_S<NEQ>(S, M.ns, M.fLL[j], A0);
  load_S<NEQ>(S, M.ns, iL, A1);
  load_S<NEQ>(S, M.ns, iR, B0);
  load_S<NEQ>(S, M.ns, M.fRR[j], B1);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const T gL = limited_regular<T>(P.limiter, P.va_eps, (T)(B0[v] - A1[v]), (T)(A0[v] - A1[v]), (T)ic);
    const T gR = limited_regular<T>(P.limiter, P.va_eps, (T)(A1[v] - B0[v]), (T)(B1[v] - B0[v]), (T)ic);
    UL[v] = A1[v] + (double)((T)pcf * gL);
    UR[v] = B0[v] + (double)((T)pcf * gR);
  }
}

/* ---- MUSCL on primitive variables (option iprimtv = 1): the reference's path of
 * deltavalues.c:16-19 (deltas of cons2primtv of both states, the finer neighbours'
 * conserved states averaged first as facevalues_MUSCL.c:31-47 does), the limiter on
 * them (reconctruct_V3.c:11-41) and prim2conc.c of the face value.  cons2primtv here
 * takes the density of the state itself and the kinetic energy from its velocities; the
 * pair keeps the reference's p = rho (gamma - 1) e - p_inf and e = (p + p_inf) /
 * (rho (gamma - 1)), which invert each other. */
template <int NEQ>
__device__ __forceinline__ void cons_to_prim(const DevParams &P, const double *U, double *W) {
  const double r = U[0];
  W[0] = r;
  W[1] = U[1] / r;
  W[2] = U[2] / r;
  W[3] = U[3] / r;
  if constexpr (NEQ == 5) {
    if (P.model == 1) {
      const double ei = (U[4] / r) - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
      W[4] = r * (P.gamma - 1.0) * ei - P.pinf;
    } else {
      W[4] = U[4] / r;
    }
  }
}

template <int NEQ>
__device__ __forceinline__ void prim_to_cons(const DevParams &P, const double *W, double *U) {
  const double r = W[0];
  U[0] = r;
  U[1] = r * W[1];
  U[2] = r * W[2];
  U[3] = r * W[3];
  if constexpr (NEQ == 5) {
    if (P.model == 1) {
      const double ei = (W[4] + P.pinf) / (r * (P.gamma - 1.0));
      U[4] = r * (ei + 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]));
    } else {
      U[4] = r * W[4];
    }
  }
}

/* recon_face on primitive variables */
template <int NEQ, typename T>
__device__ __forceinline__ void recon_face_prim(const DevMesh &M, const DevParams &P,
                                                const double *__restrict__ S, int c, int f,
                                                double *out) {
  const uint32_t cf = M.cflag[c];
  const int fo = f ^ (f < 4 ? 2 : 1);
  double U[NEQ], U0[NEQ], U1[NEQ], W[NEQ], W0[NEQ], W1[NEQ], Wf[NEQ];
  load_S<NEQ>(S, M.ns, c, U);
  nb_state<NEQ>(M, P, S, c, cf, f, U, U0);
  nb_state<NEQ>(M, P, S, c, cf, fo, U, U1);
  cons_to_prim<NEQ>(P, U, W);
  cons_to_prim<NEQ>(P, U0, W0);
  cons_to_prim<NEQ>(P, U1, W1);
  double pcf, pcc, mcf, mcc, pic, mic;
  face_dist(M, cf, f, pcf, pcc, pic);
  face_dist(M, cf, fo, mcf, mcc, mic);
  const bool bnd = ((cf >> (2 * f)) & 3) == CLS_BOUNDARY;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const T g = limited<T>(P.limiter, P.va_eps, bnd, (T)(W0[v] - W[v]), (T)(W1[v] - W[v]), (T)pic,
                           (T)mic);
    Wf[v] = W[v] + (double)((T)pcf * g);
  }
  prim_to_cons<NEQ>(P, Wf, out);
}

/* recon_regular on primitive variables */
template <int NEQ, typename T>
__device__ __forceinline__ void recon_regular_prim(const DevMesh &M, const DevParams &P,
                                                   const double *__restrict__ S, int j, int iL,
                                                   int iR, int lev, int a, double *UL, double *UR) {
  const int k = (lev * 3 + a) * 3;
  const double pcf = M.dist[2 * k], ic = M.idist[k];
  double A0[NEQ], A1[NEQ], B0[NEQ], B1[NEQ], WL[NEQ], WR[NEQ];
  load_S<NEQ>(S, M.ns, M.fLL[j], A0);
  load_S<NEQ>(S, M.ns, iL, A1);
  load_S<NEQ>(S, M.ns, iR, B0);
  load_S<NEQ>(S, M.ns, M.fRR[j], B1);
  double PA0[NEQ], PA1[NEQ], PB0[NEQ], PB1[NEQ];
  cons_to_prim<NEQ>(P, A0, PA0);
  cons_to_prim<NEQ>(P, A1, PA1);
  cons_to_prim<NEQ>(P, B0, PB0);
  cons_to_prim<NEQ>(P, B1, PB1);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const T gL = limited<T>(P.limiter, P.va_eps, false, (T)(PB0[v] - PA1[v]), (T)(PA0[v] - PA1[v]),
                            (T)ic, (T)ic);
    const T gR = limited<T>(P.limiter, P.va_eps, false, (T)(PA1[v] - PB0[v]), (T)(PB1[v] - PB0[v]),
                            (T)ic, (T)ic);
    WL[v] = PA1[v] + (double)((T)pcf * gL);
    WR[v] = PB0[v] + (double)((T)pcf * gR);
  }
  prim_to_cons<NEQ>(P, WL, UL);
  prim_to_cons<NEQ>(P, WR, UR);
}

/* ---- FP32 reconstruction of the mixed form (ideal gas): the gather keeps an
 * FP32 copy Sf of the cell-major state beside the FP64 one, and the face kernel
 * reconstructs from it (half the bytes of every gathered state). */
template <int NEQ>
__device__ __forceinline__ void load_Sf(const float *__restrict__ Sf, int i, float *U) {
  const float *p = Sf + (size_t)i * NEQ;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) U[v] = p[v];
}

template <int NEQ>
__device__ __forceinline__ void nb_state_f(const DevMesh &M, const DevParams &P,
                                           const float *__restrict__ Sf, int i, uint32_t cf, int f,
                                           const float *Uown, float *Un) {
  int cls = (cf >> (2 * f)) & 3;
  int nb = M.nbr[(size_t)f * M.ns + i];
  if (cls == CLS_SAME || cls == CLS_COARSER) {
    load_Sf<NEQ>(Sf, nb, Un);
  } else if (cls == CLS_FINER) {
    int q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = M.quad[4 * nb + k];
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc += Sf[(size_t)q[k] * NEQ + v] * 0.25f;
      Un[v] = acc;
    }
  } else {
    int bc = -1 - nb;
    if (bc == 1) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Un[v] = (float)P.inlet[v];
    } else {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Un[v] = Uown[v];
    }
  }
}

/* recon_face in FP32 throughout, from the FP32 state copy */
template <int NEQ>
__device__ __forceinline__ void recon_face_f(const DevMesh &M, const DevParams &P,
                                             const float *__restrict__ Sf, int c, int f,
                                             float *out) {
  const uint32_t cf = M.cflag[c];
  const int fo = f ^ (f < 4 ? 2 : 1);
  float U[NEQ], U0[NEQ], U1[NEQ];
  load_Sf<NEQ>(Sf, c, U);
  nb_state_f<NEQ>(M, P, Sf, c, cf, f, U, U0);
  nb_state_f<NEQ>(M, P, Sf, c, cf, fo, U, U1);
  double pcf, pcc, mcf, mcc, pic, mic;
  face_dist(M, cf, f, pcf, pcc, pic);
  face_dist(M, cf, fo, mcf, mcc, mic);
  const bool bnd = ((cf >> (2 * f)) & 3) == CLS_BOUNDARY;
  const float fp = (float)pcf, ip = (float)pic, im = (float)mic;
#pragma unroll
  for (int v = 0; v < NEQ; ++v)
    out[v] = U[v] + fp * limited<float>(P.limiter, P.va_eps, bnd, U0[v] - U[v], U1[v] - U[v], ip, im);
}

struct fvg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  fvg_params prm;
  DevParams dp;
  int neq = 5;
  bool has_mesh = false;
  int N = 0, ns = 0, F = 0, nfp = 0, Q = 0, nlevel = 0;
  int nfa[3] = {0, 0, 0};
  int nreg[3] = {0, 0, 0}; /* regular faces lead each axis group */
  /* device buffers */
  uint32_t *cflag = nullptr;
  int32_t *nbr = nullptr, *quad = nullptr, *fL = nullptr, *fR = nullptr, *gslot = nullptr;
  int32_t *fLL = nullptr, *fRR = nullptr;
  uint32_t *fmeta = nullptr;
  double *tab = nullptr;
  double *S = nullptr, *SN = nullptr, *RHS = nullptr, *SG = nullptr;
  /* spare state buffer: the time loop runs its stages out of place over S, SN and
   * Sx and swaps the three pointers at the end of each step (so S and SN keep
   * their meaning, but device addresses taken from them change) */
  double *Sx = nullptr;
  /* after fvg_steps: the buffer holding the last stage's input (Sx, or SN for a
   * one-stage method); null once anything else writes the state buffers */
  double *stage_in = nullptr;
  void *SF = nullptr, *SGm = nullptr, *G = nullptr;
  int *err = nullptr;
  double *red = nullptr;
  unsigned long long *redu = nullptr;
  size_t red_cap = 0;
  size_t rsize = 8;  /* bytes of R (8 plain, 4 mixed) */
  std::string last;
  hipEvent_t ev[2 * 64 + 2];
  bool ev_ok = false;
  /* FP32 copy of S for the mixed ideal-gas face kernel; valid while nothing but
   * that path's gather has written S since it was filled.  A stage reads Sf and writes
   * the copy of its output into Sfx, then the two swap (the LDS-tiled stage reads its
   * neighbour tiles' Sf while other workgroups write theirs). */
  float *Sf = nullptr, *Sfx = nullptr;
  bool sf_valid = false;

  /* LDS-tiled stage (fvtile.hip): the interior tiles (8^3 Morton blocks of one level
   * whose six face-neighbour blocks are such tiles too) as [ntile][8] (first cell, the
   * first cells of the neighbour tiles across faces 0..5, level); per block of the
   * stage's face kernel a flag (1: all its faces are the tiles'), per block of cells one
   * (0 none, 1 all, 2 some of its cells are the tiles') */
  int32_t *tiles = nullptr;
  int ntile = 0;
  uint8_t *face_skip = nullptr, *gather_skip = nullptr;
  uint8_t *tile_cell = nullptr; /* per cell: 1 in an interior tile (only where a cell block
                                   is partly covered: gather_skip 2) */
  bool face_all_skip = false, gather_all_skip = false;
  int tiling = FVG_TILING_AUTO; /* fvg_set_tiling */
  /* six time steps of fvg_steps captured as one HIP graph (fvg_set_graphs), replayed
   * while everything its launches bake in (pointers, sizes, parameters, dt, method) is
   * unchanged: step_graph_key holds those */
  bool graphs = false;
  hipGraphExec_t step_graph = nullptr;
  std::vector<unsigned char> step_graph_key;
  /* the skip flags of the stage being launched (null: the face kernel and the gather
   * cover the whole mesh) */
  const uint8_t *stage_fskip = nullptr, *stage_gskip = nullptr, *stage_cmask = nullptr;

  /* device address of the static flag staging array of this context's GPU (taken once
   * per device at fvg_create; fvg_adapt_flags_static) and the status of that lookup */
  int8_t *flag_scratch = nullptr;
  int flag_scratch_rc = 0;

  /* adjoint / tangent buffers (fvadj.hip), released with the mesh */
  struct fv_adj_state *adj = nullptr;

  /* leaf keys of a mesh built on the device (fvmesh.hip) and their hash table
   * (key -> cell); the next device re-flatten carries the state over by key.
   * Null for a mesh uploaded by fvg_set_mesh. */
  uint64_t *keys = nullptr, *hkeys = nullptr;
  int32_t *hval = nullptr;
  int hbits = 0;
  /* a device re-flatten between fvg_remesh_begin and fvg_remesh_end */
  struct fv_remesh *rm = nullptr;
  /* mesh blocks (fv_dmalloc / fv_dfree): handed out, and freed but kept for reuse */
  std::vector<std::pair<void *, size_t>> dlive, dcache;
};

/* Device blocks of the mesh arrays.  A rebuild after an adaptation frees and allocates a
 * few hundred MB of arrays of nearly the same sizes; hipMalloc / hipFree of them cost
 * milliseconds, so freed blocks are kept and handed out again to a request that fits
 * (at most twice its size).  Every free happens after the context's stream has drained,
 * and every later use of a block is ordered on that stream.  slack adds a quarter of
 * room to a fresh block, for meshes that grow from one rebuild to the next.  The cache
 * holds at most as many bytes as are live once a rebuild is done (fv_dcache_trim), and
 * is emptied when an allocation fails and at fvg_destroy. */
hipError_t fv_dmalloc_raw(fvg_ctx *c, void **p, size_t bytes, bool slack);
template <class T>
inline hipError_t fv_dmalloc(fvg_ctx *c, T **p, size_t bytes, bool slack) {
  void *q = nullptr;
  const hipError_t e = fv_dmalloc_raw(c, &q, bytes, slack);
  *p = (T *)q;
  return e;
}
void fv_dfree(fvg_ctx *c, void *p);
void fv_dcache_trim(fvg_ctx *c);
void fv_dcache_clear(fvg_ctx *c);

/* releases ctx->adj; defined in fvadj.hip */
void fv_adj_release(fvg_ctx *c);
/* releases ctx->rm; defined in fvmesh.hip */
void fv_remesh_release(fvg_ctx *c);
/* mesh arrays and tables (defined in fvgpu.hip, shared with fvmesh.hip) */
int fv_alloc_mesh(fvg_ctx *c, int N, int F, int Q, int nl, const int32_t nfa[3],
                  const int32_t nreg[3], bool topo);
int fv_upload_tables(fvg_ctx *c, int nl, const double *h, const double *dist, const double *area,
                     const double *vol);
/* forward pieces shared with fvadj.hip (defined in fvgpu.hip) */
void fv_launch_recon(fvg_ctx *c, const double *S);
void fv_stage(fvg_ctx *c, int irk, double dt, bool first, bool write_rhs);
/* the nine momentum gradients (gradcalc, rows 1..3; SGm[3 i + j] = d(rho u_i)/dx_j) of S
 * into c->SGm (precision of the context), for the adjoint and tangent-linear kernels (the
 * time loop stores the viscous rows of vrows there instead) */
void fv_launch_grad_raw(fvg_ctx *c, const double *S);
/* the same gradients in FP64 of a tangent state Sd (the inlet ghost, a constant, has
 * tangent zero) into SGmd [9][ns] */
void fv_launch_grad_tangent(fvg_ctx *c, const double *Sd, double *SGmd);
/* one stage Sout = RK(irk; SN, S, R(S)) without touching c->S / c->SN */
void fv_stage_ptrs(fvg_ctx *c, const double *S, double *SN, double *Sout, int irk, double dt);
void fv_launch_gather_ptrs(fvg_ctx *c, const double *G, double *S, double *SN, int irk, double dt,
                           bool first);

static inline int set_err(fvg_ctx *c, int code, const char *msg) {
  if (c) c->last = msg ? msg : "";
  return code;
}

#define HIPCHK(c, call)                                                           \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      char b_[256];                                                               \
      snprintf(b_, sizeof b_, "%s: %s", #call, hipGetErrorString(e_));            \
      return set_err((c), FVG_ERR_HIP, b_);                                       \
    }                                                                             \
  } while (0)

static inline DevMesh dev_mesh(const fvg_ctx *c) {
  DevMesh m;
  m.N = c->N;
  m.ns = c->ns;
  m.F = c->F;
  m.nfp = c->nfp;
  m.Q = c->Q;
  m.nlevel = c->nlevel;
  m.cflag = c->cflag;
  m.nbr = c->nbr;
  m.quad = c->quad;
  m.fL = c->fL;
  m.fR = c->fR;
  m.fLL = c->fLL;
  m.fRR = c->fRR;
  m.fmeta = c->fmeta;
  m.gslot = c->gslot;
  const int nl = c->nlevel;
  m.h = c->tab;
  m.dist = c->tab + 3 * nl;
  m.area = c->tab + 3 * nl + 18 * nl;
  m.vol = c->tab + 3 * nl + 18 * nl + 3 * nl;
  m.idist = c->tab + 25 * nl;
  return m;
}

/* ---- LDS-tiled stage (fvtile.hip) */
/* Find the interior tiles of a mesh uploaded by fvg_set_mesh and upload their table and
 * the skip flags of the stage's face kernel and gather; no tiles for other meshes. */
int fv_tiles_build(fvg_ctx *c, const fvg_mesh *m);
void fv_tiles_release(fvg_ctx *c);
/* One RK stage of the interior tiles' cells: MUSCL, flux, residual and the stage update
 * in LDS; S -> Sout (out of place), SN, optional RHS and FP32 copy of Sout. */
void fv_launch_tile_stage(fvg_ctx *c, const double *S, const double *SN, double *Sout, float *SfOut,
                          int irk, double dt, bool write_rhs, bool fast);
/* The momentum (and p / rho) gradients of the interior tiles' cells into SGm, bitwise
 * those of k_grad, which skips those cells' blocks while c->stage_gskip is set. */
void fv_launch_tile_grad(fvg_ctx *c, const double *S, const float *Sf = nullptr);
/* whether the tiles can run this stage (mesh, options, precision path) */
bool fv_tiles_usable(const fvg_ctx *c);

