/*
 * fvgpu.hip — gfx950 kernels and the fvg_* C ABI of the RHS + RK hot path.
 *
 * Data layout in HBM (one context = one mesh):
 *   S, SN, RHS          f64 [neq][ns]          conserved state, RK base, residual
 *   SF                  R   [6][neq][ns]       MUSCL face states (R = f64 plain, f32 mixed)
 *   SG                  f64 [neq+1][3][ns]     full Green-Gauss gradients (gradcalc)
 *   SGm                 R   [9][ns]            viscous rows (vrows) and grad(p / rho)
 *   G                   R   [neq][nfp]         owner-side face flux, -Area*(F.n + HLLC)
 *   cflag/nbr/gslot     per cell               classes, neighbours, gather slots
 *   fL/fR/fmeta         per face, grouped by axis
 * ns = ncell rounded up to 64, nfp = nface rounded up to 64.
 *
 * Kernels per RK stage (each a coalesced stream over cells or faces):
 *   k_recon      faceinterp.c:21-61     S -> SF
 *   k_gradmom    gradcalc.c:23-42       S -> SGm (only when ns == 1)
 *   k_face       residual_faces_AMR.c   SF (+SGm) -> G, each face once
 *   k_gather_rk  residual_faces_AMR.c:173-175 + RK.c:24-47
 *                                       G -> RHS (optional) -> S
 * The gather reads each cell's face fluxes in the reference's order
 * f = 0..5, k = 0..3 (no atomics, bitwise deterministic).
 */
#include <mutex>

#include "fv_common.h"

/* ------------------------------------------------------------------ kernels */

template <int MODEL, typename R, bool PRIM = false>
__global__ __launch_bounds__(FV_BLOCK) void k_recon(DevMesh M, DevParams P,
                                                    const double *__restrict__ S,
                                                    R *__restrict__ SF) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int i = logical_block() * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  if constexpr (PRIM) {
    /* MUSCL on primitive variables (option): face by face */
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
      double out[NEQ];
      recon_face_prim<NEQ, double>(M, P, S, i, f, out);
#pragma unroll
      for (int v = 0; v < NEQ; ++v) SF[((size_t)f * NEQ + v) * M.ns + i] = (R)out[v];
    }
    return;
  }
  const uint32_t cf = M.cflag[i];
  double U[NEQ];
  load_S<NEQ>(S, M.ns, i, U);
  /* axis pairs (f+, f-): x (1,3), y (2,0), z (5,4) */
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int fp = ax == 0 ? 1 : (ax == 1 ? 2 : 5);
    const int fm = ax == 0 ? 3 : (ax == 1 ? 0 : 4);
    double Up[NEQ], Um[NEQ];
    nb_state<NEQ>(M, P, S, i, cf, fp, U, Up);
    nb_state<NEQ>(M, P, S, i, cf, fm, U, Um);
    double pcf, pcc, mcf, mcc, pic, mic;
    face_dist(M, cf, fp, pcf, pcc, pic);
    face_dist(M, cf, fm, mcf, mcc, mic);
    const bool bp = ((cf >> (2 * fp)) & 3) == CLS_BOUNDARY;
    const bool bm = ((cf >> (2 * fm)) & 3) == CLS_BOUNDARY;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      double dP = Up[v] - U[v];
      double dM = Um[v] - U[v];
      double gp = limited<double>(P.limiter, P.va_eps, bp, dP, dM, pic, mic);
      double gm = limited<double>(P.limiter, P.va_eps, bm, dM, dP, mic, pic);
      SF[((size_t)fp * NEQ + v) * M.ns + i] = (R)(U[v] + pcf * gp);
      SF[((size_t)fm * NEQ + v) * M.ns + i] = (R)(U[v] + mcf * gm);
    }
  }
}

/* Green-Gauss gradients (gradcalc.c:23-42).  FULL: all neq+1 rows (incl. p) to
 * SG f64 [neq+1][3][ns]; otherwise the gradients of the momenta (rows 1..3) as the
 * six viscous rows of vrows to SGm [9][ns] (RAW: the nine gradients themselves,
 * SGm[3 i + j] = d(rho u_i)/dx_j, for the adjoint and tangent-linear kernels).
 * SS: the state's storage, the FP64 state or (the mixed form's viscous rows, which
 * feed its FP32 viscous flux) its FP32 copy; the sums are FP64 either way. */
template <int MODEL, bool FULL, typename R, bool HEAT = false, bool RAW = false, typename SS = double>
__global__ __launch_bounds__(FV_BLOCK) void k_grad(DevMesh M, DevParams P,
                                                   const SS *__restrict__ S,
                                                   double *__restrict__ SG, R *__restrict__ SGm,
                                                   const uint8_t *__restrict__ skip,
                                                   const uint8_t *__restrict__ cmask) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  /* HEAT (heat conduction option, model 1): also the Green-Gauss gradient of p / rho,
   * from the same face means, into SGm rows 9..11 */
  static_assert(!HEAT || (MODEL == 1 && !FULL && !RAW), "heat conduction: model 1, viscous rows");
  const int lb = logical_block();
  /* a block of cells that the LDS-tiled gradient kernel covers exits whole; in a block it
   * covers in part, its cells do */
  const int sk = skip ? skip[lb] : 0;
  if (sk == 1) return;
  const int i = lb * FV_BLOCK + threadIdx.x;
  if (i >= M.N || (sk == 2 && cmask[i])) return;
  const uint32_t cf = M.cflag[i];
  const int L = (cf >> 16) & 0xff;
  const double V = M.vol[L];
  double U[NEQ];
  load_S<NEQ>(S, M.ns, i, U);
  double pown = 0.0, town = 0.0;
  if constexpr (FULL) {
    double r, u[3], e, p, c, y;
    eos64<MODEL>(P, U, false, r, u, e, p, c, y);
    pown = p;
  }
  if constexpr (HEAT) town = theta_ideal(P, U);
  constexpr int V0 = FULL ? 0 : 1;
  constexpr int V1 = FULL ? NEQ : 4;
  /* the six neighbours and the row of each single neighbour are loaded before any
   * branch (the cell's own row stands in for the other classes), so that all of them
   * are in flight together; the sums below keep their order */
  int nbs[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) nbs[f] = M.nbr[(size_t)f * M.ns + i];
  double Ur[6][NEQ];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int cls = (cf >> (2 * f)) & 3;
    load_S<NEQ>(S, M.ns, (cls == CLS_SAME || cls == CLS_COARSER) ? nbs[f] : i, Ur[f]);
  }
  /* theta of the single neighbours, formed right away so that only the momentum rows
   * of Ur stay live */
  double thr[6];
  if constexpr (HEAT) {
#pragma unroll
    for (int f = 0; f < 6; ++f) thr[f] = theta_ideal(P, Ur[f]);
  }
  double gm[9]; /* gm[3 i + d] = d(rho u_i)/dx_d */
  /* per axis d, faces of that axis in increasing face id */
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int fa = d == 0 ? 1 : (d == 1 ? 0 : 4);
    const int fb = d == 0 ? 3 : (d == 1 ? 2 : 5);
    double acc[NEQ + 1], acct = 0.0;
#pragma unroll
    for (int v = 0; v <= NEQ; ++v) acc[v] = 0.0;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int f = t == 0 ? fa : fb;
      const double nd = (double)c_face_sign[f];
      const int cls = (cf >> (2 * f)) & 3;
      const int nb = nbs[f];
      const int nk = cls == CLS_FINER ? 4 : 1;
      const double A = M.area[3 * (cls == CLS_FINER ? L + 1 : L) + d];
      /* the time loop's rows: n A / V once per face, not one division per term (the
       * reference's term order, 0.5 (Un + U) n A / V, is kept for gradcalc()) */
      const double wgt = nd * A / V;
      for (int k = 0; k < nk; ++k) {
        double Un[NEQ];
        if (cls == CLS_FINER) load_S<NEQ>(S, M.ns, M.quad[4 * nb + k], Un);
        else if (cls == CLS_BOUNDARY) {
          if (-1 - nb == 1) {
#pragma unroll
            for (int v = 0; v < NEQ; ++v) Un[v] = P.inlet[v];
          } else {
#pragma unroll
            for (int v = 0; v < NEQ; ++v) Un[v] = U[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < NEQ; ++v) Un[v] = Ur[f][v];
        }
#pragma unroll
        for (int v = V0; v < V1; ++v) {
          if constexpr (FULL) acc[v] += 0.5 * (Un[v] + U[v]) * nd * A / V;
          else acc[v] += 0.5 * (Un[v] + U[v]) * wgt;
        }
        if constexpr (FULL) {
          double r, u[3], e, p, c, y;
          eos64<MODEL>(P, Un, false, r, u, e, p, c, y);
          acc[NEQ] += 0.5 * (p + pown) * nd * A / V;
        }
        if constexpr (HEAT) {
          const double thn = cls == CLS_FINER ? theta_ideal(P, Un)
                             : (cls == CLS_BOUNDARY ? (-1 - nb == 1 ? theta_ideal(P, Un) : town)
                                                    : thr[f]);
          acct += 0.5 * (thn + town) * wgt;
        }
      }
    }
    if constexpr (HEAT) SGm[((size_t)VROW_HEAT + d) * M.ns + i] = (R)acct;
    if constexpr (FULL) {
#pragma unroll
      for (int v = 0; v <= NEQ; ++v) SG[((size_t)v * 3 + d) * M.ns + i] = acc[v];
    } else {
#pragma unroll
      for (int v = 1; v < 4; ++v) gm[(v - 1) * 3 + d] = acc[v];
    }
  }
  if constexpr (!FULL) {
    if constexpr (RAW) {
#pragma unroll
      for (int k = 0; k < 9; ++k) SGm[(size_t)k * M.ns + i] = (R)gm[k];
    } else {
      double vr[6];
      vrows<double>(gm, vr);
#pragma unroll
      for (int k = 0; k < 6; ++k) SGm[(size_t)k * M.ns + i] = (R)vr[k];
    }
  }
}

/* Flux rows whose size is not a multiple of 16 bytes (5 doubles) are staged in LDS
 * (one row per thread) and then written as the block's contiguous range
 * [j0, j0 + nrow) of G with consecutive words per lane: full-line stores instead of
 * NEQ strided partial ones per thread.  Rows of 4 doubles go out as aligned 16-byte
 * pairs and FP32 rows are small: both measured faster stored directly. */
template <int NEQ, typename R>
struct RowStage {
  static constexpr bool on = sizeof(R) == 8 && (NEQ * sizeof(R)) % 16 != 0;
};

template <int NEQ, typename R>
__device__ __forceinline__ void store_rows(R *__restrict__ G, const R *gbuf, int j0, int nrow) {
  R *g = G + (size_t)j0 * NEQ;
  for (int k = threadIdx.x; k < nrow * NEQ; k += FV_BLOCK) g[k] = gbuf[k];
}

/* the gradients of p / rho of both cells (SGm rows 6..8; the kernels with the
 * heat-conduction bit read them, and use them where P.kheat != 0) */
template <typename T, typename R>
__device__ __forceinline__ void load_heat_grad(const DevMesh &M, const R *__restrict__ SGm,
                                               int iL, int iR, FState<T> &L, FState<T> &Rs) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    L.gt[d] = (T)SGm[(size_t)(VROW_HEAT + d) * M.ns + iL];
    Rs.gt[d] = (T)SGm[(size_t)(VROW_HEAT + d) * M.ns + iR];
  }
}

/* both cells' viscous rows of axis a (SGm rows vrow(a, 0..2)) */
template <typename T, typename R>
__device__ __forceinline__ void load_visc_rows(const DevMesh &M, const R *__restrict__ SGm, int a,
                                               int iL, int iR, FState<T> &L, FState<T> &Rs) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    L.tv[k] = (T)SGm[(size_t)vrow(a, k) * M.ns + iL];
    Rs.tv[k] = (T)SGm[(size_t)vrow(a, k) * M.ns + iR];
  }
}

/* one face, both sides (residual_faces_AMR.c:95-150 for the owner side), with
 * the axis a of the face as a compile-time constant.  OPT: the option variant (PVRS
 * wave speeds, heat conduction; the primitive reconstruction is k_recon's) */
template <int MODEL, bool NS, int a, typename T, typename R, int OPT = 0>
__device__ __forceinline__ void face_body(const DevMesh &M, const DevParams &P, int j,
                                          const double *__restrict__ S, const R *__restrict__ SF,
                                          const R *__restrict__ SGm, R *dst,
                                          int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int iL = M.fL[j], iR = M.fR[j];
  const uint32_t meta = M.fmeta[j];
  const int fl = meta & 7, fr = (meta >> 3) & 7, bc = (meta >> 6) & 3, lev = (meta >> 8) & 0xff;
  const T s = (T)c_face_sign[fl];
  double UL[NEQ], UR[NEQ];
  const bool first = P.ischeme == 1;
#pragma unroll
  for (int v = 0; v < NEQ; ++v)
    UL[v] = first ? S[(size_t)iL * NEQ + v] : (double)SF[((size_t)fl * NEQ + v) * M.ns + iL];
  if (bc == 0) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v)
      UR[v] = first ? S[(size_t)iR * NEQ + v] : (double)SF[((size_t)fr * NEQ + v) * M.ns + iR];
  } else if (bc == 1) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) UR[v] = P.inlet[v];
  } else {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) UR[v] = UL[v];
  }
  FState<T> L, Rs;
  const bool clip = !first;
  bool ok = make_state<MODEL, T>(P, UL, clip, L);
  ok = make_state<MODEL, T>(P, UR, clip && bc != 1, Rs) && ok;
  if constexpr (NS) {
    load_visc_rows(M, SGm, a, iL, iR, L, Rs);
    if constexpr ((OPT & OPT_HEAT) != 0) load_heat_grad<T, R>(M, SGm, iL, iR, L, Rs);
  }
  T out[NEQ];
  ok = hllc_face<MODEL, NS, a, T, OPT>(P, L, Rs, s, out) && ok;
  const T A = (T)M.area[3 * lev + a];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    T g = -(A) * out[v];
    dst[v] = (R)g;
  }
  if (!ok) atomicMin(err, iL);
}

/* face flux with the reconstruction fused in: reads S (the 2-ring through the
 * caches) instead of a stored SF, same results as recon + face.  REG: a regular
 * face (interior, stencil face_LL, face_L, face_R, face_RR of one level along
 * the axis): the same arithmetic as recon_face with the neighbours and the
 * same-level distances known, without the class and neighbour lookups. */
template <int MODEL, bool NS, int a, typename T, typename R, bool REG, int OPT = 0>
__device__ __forceinline__ void face_fused_body(const DevMesh &M, const DevParams &P, int j,
                                                const double *__restrict__ S,
                                                const R *__restrict__ SGm, R *dst,
                                                int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int iL = M.fL[j], iR = M.fR[j];
  const uint32_t meta = M.fmeta[j];
  const int fl = meta & 7, fr = (meta >> 3) & 7, bc = REG ? 0 : (meta >> 6) & 3,
            lev = (meta >> 8) & 0xff;
  const T s = (T)c_face_sign[fl];
  double UL[NEQ], UR[NEQ];
  const bool prim = (OPT & OPT_PRIM) && P.iprimtv;
  if constexpr (REG) {
    if (prim) recon_regular_prim<NEQ, T>(M, P, S, j, iL, iR, lev, a, UL, UR);
    else recon_regular<NEQ, T>(M, P, S, j, iL, iR, lev, a, UL, UR);
  } else {
    if (prim) recon_face_prim<NEQ, T>(M, P, S, iL, fl, UL);
    else recon_face<NEQ, T>(M, P, S, iL, fl, UL);
    if (bc == 0) {
      if (prim) recon_face_prim<NEQ, T>(M, P, S, iR, fr, UR);
      else recon_face<NEQ, T>(M, P, S, iR, fr, UR);
    } else if (bc == 1) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) UR[v] = P.inlet[v];
    } else {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) UR[v] = UL[v];
    }
  }
  FState<T> L, Rs;
  bool ok = make_state<MODEL, T>(P, UL, true, L);
  ok = make_state<MODEL, T>(P, UR, bc != 1, Rs) && ok;
  if constexpr (NS) {
    load_visc_rows(M, SGm, a, iL, iR, L, Rs);
    if constexpr ((OPT & OPT_HEAT) != 0) load_heat_grad<T, R>(M, SGm, iL, iR, L, Rs);
  }
  T out[NEQ];
  ok = hllc_face<MODEL, NS, a, T, OPT>(P, L, Rs, s, out) && ok;
  const T A = (T)M.area[3 * lev + a];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) dst[v] = (R)(-(A) * out[v]);
  if (!ok) atomicMin(err, iL);
}

template <int MODEL, bool NS, typename T, typename R, int OPT = 0>
/* occupancy targets: 6 waves per SIMD for the ideal-gas Euler flux (78-80 VGPRs, no
 * spill; at 7 waves it spills 22 VGPRs and runs 3% slower), 7 for the barotropic one
 * (72 VGPRs, 2 spilled, 1% faster than at 6); the viscous variants need 96-98 VGPRs
 * and ask for fewer waves instead of spilling.  OPT: option bits (primitive MUSCL,
 * PVRS, heat conduction). */
__global__ __launch_bounds__(FV_BLOCK, NS ? (sizeof(T) == 4 ? 5 : 4) : (OPT ? 4 : (MODEL == 2 ? 7 : 6))) void k_face_fused(
    DevMesh M, DevParams P, int n0, int n1, int n2, int r0, int r1, int r2,
    const double *__restrict__ S, const R *__restrict__ SGm, R *__restrict__ G,
    int *__restrict__ err, const uint8_t *__restrict__ skip) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  constexpr bool STAGE = RowStage<NEQ, R>::on;
  __shared__ R gbuf[STAGE ? FV_BLOCK * NEQ : 1];
  const int t = threadIdx.x;
  /* axis a: faces [0, r_a) of the group are regular (a block straddles the end of
   * that range at most once per axis) */
  int blk;
  const int ax = face_block(n0, n1, n2, r0, r1, r2, blk);
  if (skipped_face_block(skip, n0, n1, ax, blk)) return;
  const int q = blk * FV_BLOCK + t;
  const int jq = ax == 0 ? q : (ax == 1 ? n0 + q : n0 + n1 + q);
  R *dst = STAGE ? gbuf + t * NEQ : G + (size_t)jq * NEQ;
  int base, n;
  if (ax == 0) {
    base = 0;
    n = n0;
    if (q < r0) face_fused_body<MODEL, NS, 0, T, R, true, OPT>(M, P, q, S, SGm, dst, err);
    else if (q < n0) face_fused_body<MODEL, NS, 0, T, R, false, OPT>(M, P, q, S, SGm, dst, err);
  } else if (ax == 1) {
    base = n0;
    n = n1;
    if (q < r1) face_fused_body<MODEL, NS, 1, T, R, true, OPT>(M, P, n0 + q, S, SGm, dst, err);
    else if (q < n1) face_fused_body<MODEL, NS, 1, T, R, false, OPT>(M, P, n0 + q, S, SGm, dst, err);
  } else {
    base = n0 + n1;
    n = n2;
    if (q < r2) face_fused_body<MODEL, NS, 2, T, R, true, OPT>(M, P, n0 + n1 + q, S, SGm, dst, err);
    else if (q < n2) face_fused_body<MODEL, NS, 2, T, R, false, OPT>(M, P, n0 + n1 + q, S, SGm, dst, err);
  }
  if constexpr (STAGE) {
    __syncthreads();
    store_rows<NEQ, R>(G, gbuf, base + blk * FV_BLOCK, min(FV_BLOCK, n - blk * FV_BLOCK));
  }
}

/* Mixed form of the ideal gas (model 1): MUSCL from the FP32 state copy, the
 * FP32 inversion (make_state_mixed_ideal), FP32 HLLC and viscous flux, FP32
 * face flux; FP64 stays in the gather, the RK update and the time step.  A face
 * with a state the FP32 inversion rejects (near vacuum, internal energy a small
 * difference of large terms) is left to face_rebuild_block, which rebuilds it
 * from the FP64 cell states with the FP64 inversion, exactly as the FP64-read
 * mixed kernel computes it.  Returns false for such a face. */
template <bool NS, int a, bool REG, bool HEAT = false>
__device__ __forceinline__ bool face_fast_body(const DevMesh &M, const DevParams &P, int j,
                                               const float *__restrict__ Sf,
                                               const float *__restrict__ SGm,
                                               float *dst, int *__restrict__ err) {
  const int iL = M.fL[j], iR = M.fR[j];
  const uint32_t meta = M.fmeta[j];
  const int fl = meta & 7, fr = (meta >> 3) & 7, bc = REG ? 0 : (meta >> 6) & 3,
            lev = (meta >> 8) & 0xff;
  const float s = (float)c_face_sign[fl];
  FState<float> L, Rs;
  float U[5];
  bool fp32;
  if constexpr (REG) {
    /* regular face: recon_face_f with the four-cell stencil and the same-level
     * distances known (same arithmetic) */
    const int k = (lev * 3 + a) * 3;
    const float fp = (float)M.dist[2 * k], ic = (float)M.idist[k], fpic = fp * ic;
    float A0[5], A1[5], B0[5], B1[5], V[5];
    load_Sf<5>(Sf, M.fLL[j], A0);
    load_Sf<5>(Sf, iL, A1);
    load_Sf<5>(Sf, iR, B0);
    load_Sf<5>(Sf, M.fRR[j], B1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      U[v] = muscl_regular_f(P.limiter, P.va_eps, A1[v], B0[v], A0[v], fp, ic, fpic);
      V[v] = muscl_regular_f(P.limiter, P.va_eps, B0[v], A1[v], B1[v], fp, ic, fpic);
    }
    fp32 = make_state_mixed_ideal(P, U, L);
    fp32 = make_state_mixed_ideal(P, V, Rs) && fp32;
  } else {
    recon_face_f<5>(M, P, Sf, iL, fl, U);
    fp32 = make_state_mixed_ideal(P, U, L);
    if (bc == 0) {
      recon_face_f<5>(M, P, Sf, iR, fr, U);
      fp32 = make_state_mixed_ideal(P, U, Rs) && fp32;
    } else if (bc == 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) U[v] = (float)P.inlet[v];
      fp32 = make_state_mixed_ideal(P, U, Rs) && fp32;
    } else {
      Rs = L;
    }
  }
  if (!fp32) return false;
  if constexpr (NS) {
    load_visc_rows(M, SGm, a, iL, iR, L, Rs);
    if constexpr (HEAT) load_heat_grad<float, float>(M, SGm, iL, iR, L, Rs);
  }
  float out[5];
  const bool ok = hllc_face<1, NS, a, float, HEAT ? OPT_HEAT : 0>(P, L, Rs, s, out);
  const float A = (float)M.area[3 * lev + a];
#pragma unroll
  for (int v = 0; v < 5; ++v) dst[v] = -(A)*out[v];
  if (!ok) atomicMin(err, iL);
  return true;
}

/* variable v of the MUSCL neighbour state (nb_state for one variable) */
__device__ __forceinline__ double nb_var(const DevMesh &M, const DevParams &P,
                                         const double *__restrict__ S, int i, uint32_t cf, int f,
                                         int v, double own) {
  const int cls = (cf >> (2 * f)) & 3;
  const int nb = M.nbr[(size_t)f * M.ns + i];
  if (cls == CLS_SAME || cls == CLS_COARSER) return S[(size_t)nb * 5 + v];
  if (cls == CLS_FINER) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc += S[(size_t)M.quad[4 * nb + k] * 5 + v] / ((double)4);
    return acc;
  }
  return (-1 - nb) == 1 ? P.inlet[v] : own;
}

/* variable v of recon_face<5, float> (the same arithmetic, one variable) */
__device__ __forceinline__ double recon_face_var(const DevMesh &M, const DevParams &P,
                                                 const double *__restrict__ S, int c, int f, int v) {
  const uint32_t cf = M.cflag[c];
  const int fo = f ^ (f < 4 ? 2 : 1);
  const double u = S[(size_t)c * 5 + v];
  const double u0 = nb_var(M, P, S, c, cf, f, v, u), u1 = nb_var(M, P, S, c, cf, fo, v, u);
  double pcf, pcc, mcf, mcc, pic, mic;
  face_dist(M, cf, f, pcf, pcc, pic);
  face_dist(M, cf, fo, mcf, mcc, mic);
  const bool bnd = ((cf >> (2 * f)) & 3) == CLS_BOUNDARY;
  const float g = limited<float>(P.limiter, P.va_eps, bnd, (float)(u0 - u), (float)(u1 - u),
                                 (float)pic, (float)mic);
  return u + (double)((float)pcf * g);
}

/* The faces of this block that the FP32 inversion rejected (list[0 .. n)),
 * rebuilt as the FP64-read mixed kernel computes them (MUSCL from the FP64
 * states with the FP32 limiter, FP64 inversion, FP32 flux).  25 faces at a time:
 * one thread per (face, side, variable) stages the FP64 face states in LDS, then
 * one thread per face inverts them and evaluates the flux.  Written lean so the
 * kernel's register budget stays close to that of its main path. */
template <bool NS, bool HEAT = false>
__device__ void face_rebuild_block(const DevMesh &M, const DevParams &P,
                                   const double *__restrict__ S, const float *__restrict__ SGm,
                                   float *__restrict__ G, int *__restrict__ err, const int *list,
                                   int n) {
  constexpr int CH = 25;
  __shared__ double sX[CH][2][5];
  const int t = threadIdx.x;
  for (int base = 0; base < n; base += CH) {
    if (t < CH * 10) {
      const int fi = t / 10, side = (t / 5) & 1, v = t % 5;
      if (base + fi < n) {
        const int j = list[base + fi];
        const uint32_t meta = M.fmeta[j];
        const int fl = meta & 7, fr = (meta >> 3) & 7, bc = (meta >> 6) & 3;
        double x;
        if (side == 0 || bc == 2) x = recon_face_var(M, P, S, M.fL[j], fl, v);
        else if (bc == 0) x = recon_face_var(M, P, S, M.fR[j], fr, v);
        else x = P.inlet[v];
        sX[fi][side][v] = x;
      }
    }
    __syncthreads();
    if (t < CH && base + t < n) {
      const int j = list[base + t];
      const uint32_t meta = M.fmeta[j];
      const int fl = meta & 7, bc = (meta >> 6) & 3, lev = (meta >> 8) & 0xff;
      const int iL = M.fL[j], iR = M.fR[j];
      const int a = c_face_axis[fl];
      FState<float> L, Rs;
      bool ok = make_state<1, float>(P, sX[t][0], true, L);
      ok = make_state<1, float>(P, sX[t][1], bc != 1, Rs) && ok;
      if constexpr (NS) {
        load_visc_rows(M, SGm, a, iL, iR, L, Rs);
        if constexpr (HEAT) load_heat_grad<float, float>(M, SGm, iL, iR, L, Rs);
      }
      const float s = (float)c_face_sign[fl];
      float out[5];
      constexpr int O = HEAT ? OPT_HEAT : 0;
      if (a == 0) ok = hllc_face<1, NS, 0, float, O>(P, L, Rs, s, out) && ok;
      else if (a == 1) ok = hllc_face<1, NS, 1, float, O>(P, L, Rs, s, out) && ok;
      else ok = hllc_face<1, NS, 2, float, O>(P, L, Rs, s, out) && ok;
      const float A = (float)M.area[3 * lev + a];
      for (int v = 0; v < 5; ++v) G[(size_t)j * 5 + v] = -(A)*out[v];
      if (!ok) atomicMin(err, iL);
    }
    __syncthreads();
  }
}

/* The faces of a block whose FP32 inversion was rejected are collected in LDS
 * and rebuilt by the block itself (no other block, no extra launch); blocks with
 * no such face skip it after one block-wide vote. */
template <bool NS, bool HEAT = false>
__global__ __launch_bounds__(FV_BLOCK, NS ? 7 : 8) void k_face_fast(DevMesh M, DevParams P, int n0, int n1, int n2,
                                                                   int r0, int r1, int r2,
                                                                   const double *__restrict__ S,
                                                                   const float *__restrict__ Sf,
                                                                   const float *__restrict__ SGm,
                                                                   float *__restrict__ G,
                                                                   int *__restrict__ err,
                                                                   const uint8_t *__restrict__ skip) {
  __shared__ int list[FV_BLOCK];
  __shared__ int nlist;
  const int t = threadIdx.x;
  int blk;
  const int ax = face_block(n0, n1, n2, r0, r1, r2, blk);
  if (skipped_face_block(skip, n0, n1, ax, blk)) return;
  if (t == 0) nlist = 0;
  __syncthreads();
  bool done = true;
  const int q = blk * FV_BLOCK + t;
  /* FP32 rows (20 B) are stored directly: staging them through LDS as the FP64
   * kernels do measured slower here */
  const int j = ax == 0 ? q : (ax == 1 ? n0 + q : n0 + n1 + q);
  float *dst = G + (size_t)j * 5;
  if (ax == 0) {
    if (q < r0) done = face_fast_body<NS, 0, true, HEAT>(M, P, j, Sf, SGm, dst, err);
    else if (q < n0) done = face_fast_body<NS, 0, false, HEAT>(M, P, j, Sf, SGm, dst, err);
  } else if (ax == 1) {
    if (q < r1) done = face_fast_body<NS, 1, true, HEAT>(M, P, j, Sf, SGm, dst, err);
    else if (q < n1) done = face_fast_body<NS, 1, false, HEAT>(M, P, j, Sf, SGm, dst, err);
  } else {
    if (q < r2) done = face_fast_body<NS, 2, true, HEAT>(M, P, j, Sf, SGm, dst, err);
    else if (q < n2) done = face_fast_body<NS, 2, false, HEAT>(M, P, j, Sf, SGm, dst, err);
  }
  if (!__syncthreads_or(!done)) return;
  if (!done) list[atomicAdd(&nlist, 1)] = j;
  __syncthreads();
  face_rebuild_block<NS, HEAT>(M, P, S, SGm, G, err, list, nlist);
}

/* Refinement flags from the density gradient SG[0][.] (gradcalc rows, [neq+1][3][ns]):
 * kind 0: g = h |grad rho| / rho (scale-free), kind 1: g = |grad rho|, the quantity of
 * criterionsplt.c:60-69.  +1 (split) where g > refine and the level is below
 * max_level, -1 (merge) where g < coarsen and the level is above base_level. */
__global__ void k_adapt_flags(DevMesh M, const double *__restrict__ S, int neq,
                              const double *__restrict__ SG, int kind, double refine,
                              double coarsen, int base_level, int max_level,
                              int8_t *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.N) return;
  const int L = (M.cflag[i] >> 16) & 0xff;
  const double gx = SG[(size_t)0 * M.ns + i], gy = SG[(size_t)1 * M.ns + i],
               gz = SG[(size_t)2 * M.ns + i];
  double g = sqrt(gx * gx + gy * gy + gz * gz);
  if (kind == 0) g = g * M.h[3 * L] / S[(size_t)i * neq];
  int8_t f = 0;
  if (g > refine && L < max_level) f = 1;
  else if (g < coarsen && L > base_level) f = -1;
  flags[i] = f;
}

/* cell pressure with the gradient kernel's FP64 EOS (the p row of gradcalc) */
template <int MODEL>
__global__ void k_pressure(DevMesh M, DevParams P, const double *__restrict__ S,
                           double *__restrict__ pout) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.N) return;
  double U[NEQ];
  load_S<NEQ>(S, M.ns, i, U);
  double r, u[3], e, p, c, y;
  eos64<MODEL>(P, U, false, r, u, e, p, c, y);
  pout[i] = p;
}

/* S (FP64, cell-major) -> its FP32 copy */
__global__ void k_to_f32(const double *__restrict__ S, float *__restrict__ Sf, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) Sf[k] = (float)S[k];
}

/* Faces are grouped by axis; blocks [0, b0) take x faces, [b0, b0+b1) y faces,
 * the rest z faces, so the axis is uniform within a block. */
template <int MODEL, bool NS, typename T, typename R, int OPT = 0>
__global__ __launch_bounds__(FV_BLOCK) void k_face(DevMesh M, DevParams P, int b0, int b1,
                                                   int n0, int n1, int n2,
                                                   const double *__restrict__ S,
                                                   const R *__restrict__ SF,
                                                   const R *__restrict__ SGm,
                                                   R *__restrict__ G, int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  constexpr bool STAGE = RowStage<NEQ, R>::on;
  __shared__ R gbuf[STAGE ? FV_BLOCK * NEQ : 1];
  const int lb = logical_block();
  const int t = threadIdx.x;
  int blk;
  const int ax = axis_chunk(lb, b0, b1, blk);
  const int q = blk * FV_BLOCK + t;
  const int base = ax == 0 ? 0 : (ax == 1 ? n0 : n0 + n1);
  const int n = ax == 0 ? n0 : (ax == 1 ? n1 : n2);
  R *dst = STAGE ? gbuf + t * NEQ : G + (size_t)(base + q) * NEQ;
  if (ax == 0) {
    if (q < n0) face_body<MODEL, NS, 0, T, R, OPT>(M, P, q, S, SF, SGm, dst, err);
  } else if (ax == 1) {
    if (q < n1) face_body<MODEL, NS, 1, T, R, OPT>(M, P, n0 + q, S, SF, SGm, dst, err);
  } else {
    if (q < n2) face_body<MODEL, NS, 2, T, R, OPT>(M, P, n0 + n1 + q, S, SF, SGm, dst, err);
  }
  if constexpr (STAGE) {
    __syncthreads();
    store_rows<NEQ, R>(G, gbuf, base + blk * FV_BLOCK, min(FV_BLOCK, n - blk * FV_BLOCK));
  }
}

/* RHS gather (residual_faces_AMR.c:173-175: the sum over a cell's faces and their
 * sub-faces k = 0..3, here in the face order -x, +x, -y, +y, -z, +z where the reference
 * runs f = 0..5) fused with the RK stage (RK.c:24-47).  irk 1..4; FIRST: also SN = S
 * (RK(0)).
 * A block owns 256 consecutive cells, whose cell-major S / SN / RHS rows form one
 * contiguous span: it is moved between HBM and LDS with fully coalesced accesses
 * and each thread works on its own row in LDS.  The stage reads S and writes Sout,
 * which is S itself in the time loop and another buffer when the adjoint
 * recomputes stage states into its segment store (each block reads its whole span
 * before any thread writes, so Sout == S is safe).  Sf (optional) receives the
 * FP32 copy of the new state for the mixed ideal-gas face kernel. */
template <int NEQ, typename R, bool WRITE_RHS, bool DO_RK, bool FIRST>
__global__ __launch_bounds__(FV_BLOCK) void k_gather_rk(DevMesh M, const R *__restrict__ G,
                                                        const double *S,
                                                        double *SN,
                                                        double *__restrict__ RHS, int irk,
                                                        double dt, int *__restrict__ err,
                                                        double *Sout, float *Sf,
                                                        const uint8_t *__restrict__ skip,
                                                        const uint8_t *__restrict__ cmask) {
  __shared__ double shS[FV_BLOCK * NEQ];
  __shared__ double shN[FV_BLOCK * NEQ];
  const int lb = logical_block();
  const int t = threadIdx.x;
  const int i0 = lb * FV_BLOCK;
  /* the grid is padded to a multiple of 8 blocks (gather_grid): the padding exits whole;
   * so does a block whose cells the LDS-tiled stage updates.  In a block whose cells it
   * updates in part (skip 2), cm marks them: their rows are neither checked nor written. */
  const int sk = (skip && i0 < M.N) ? skip[lb] : 0;
  if (i0 >= M.N || sk == 1) return;
  const uint8_t *cm = sk == 2 ? cmask : nullptr;
  const int i = i0 + t;
  const int nrow = min(FV_BLOCK, M.N - i0);
  const int nval = nrow * NEQ;
  const size_t base = (size_t)i0 * NEQ;
  /* The span's S / SN values are loaded into registers first and written to LDS only
   * after the flux loads below have been issued, so that the two chains of loads
   * overlap (on small meshes the stage has too few waves to hide one latency behind
   * another).  The first stage of the out-of-place time step has SN == S: read it
   * once. */
  const bool one = FIRST || SN == S;
  double ls[NEQ], ln[NEQ];
  if constexpr (DO_RK) {
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
      const int k = t + q * FV_BLOCK;
      if (k < nval) {
        ls[q] = S[base + k];
        if (!one) ln[q] = SN[base + k];
      }
    }
  }
  double rhs[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) rhs[v] = 0.0;
  const bool live = i < M.N;
  uint32_t cf = 0;
  if (live) {
    cf = M.cflag[i];
    /* the six slots and the first flux row of each face are loaded up front, outside
     * any branch, so that the loads of all faces are in flight together (a load
     * under a per-face branch is waited for before the next face's is issued); a
     * face without a slot reads row 0 and ignores it; only the three further rows
     * of a finer face are loaded in its branch */
    int g[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) g[f] = M.gslot[(size_t)f * M.ns + i];
    R r0[6][NEQ];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const size_t j0 = g[f] >= 0 ? (size_t)(g[f] >> 1) : 0;
#pragma unroll
      for (int v = 0; v < NEQ; ++v) r0[f][v] = G[j0 * NEQ + v];
    }
    /* axis by axis: -x, +x, -y, +y, -z, +z (the LDS-tiled stage sums a cell's faces in
     * the same order, one axis per pass) */
    constexpr int ford[6] = {3, 1, 0, 2, 4, 5};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int f = ford[q];
      const int cls = (cf >> (2 * f)) & 3;
      double ff[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; ++v) ff[v] = 0.0;
      if (g[f] >= 0) {
        const int jf = g[f] >> 1;
        if (cls == CLS_FINER) {
#pragma unroll
          for (int v = 0; v < NEQ; ++v) ff[v] += (double)r0[f][v];
#pragma unroll
          for (int k = 1; k < 4; ++k)
#pragma unroll
            for (int v = 0; v < NEQ; ++v) ff[v] += (double)G[(size_t)(jf + k) * NEQ + v];
        } else if (g[f] & 1) {
#pragma unroll
          for (int v = 0; v < NEQ; ++v) ff[v] += -(double)r0[f][v];
        } else {
#pragma unroll
          for (int v = 0; v < NEQ; ++v) ff[v] += (double)r0[f][v];
        }
      }
#pragma unroll
      for (int v = 0; v < NEQ; ++v) rhs[v] += ff[v];
    }
    bool bad = false;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) bad |= (rhs[v] != rhs[v]);
    if (bad && !(cm && cm[i])) atomicMin(err, i);
  }
  if constexpr (DO_RK) {
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
      const int k = t + q * FV_BLOCK;
      if (k < nval) {
        shS[k] = ls[q];
        if constexpr (!FIRST) shN[k] = one ? ls[q] : ln[q];
      }
    }
  }
  __syncthreads();
  if (live) {
    if constexpr (DO_RK) {
      const double Vol = M.vol[(cf >> 16) & 0xff];
#pragma unroll
      for (int v = 0; v < NEQ; ++v) {
        const int o = t * NEQ + v;
        const double s = shS[o];
        double sn;
        if constexpr (FIRST) { sn = s; shN[o] = s; } else { sn = shN[o]; }
        shS[o] = rk_update(irk, s, sn, rhs[v], dt, Vol);
      }
    } else if constexpr (WRITE_RHS) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) shS[t * NEQ + v] = rhs[v];
    }
  }
  __syncthreads();
  if constexpr (DO_RK) {
    for (int k = t; k < nval; k += FV_BLOCK) {
      if (cm && cm[i0 + k / NEQ]) continue;
      Sout[base + k] = shS[k];
      if (Sf) Sf[base + k] = (float)shS[k];
      if constexpr (FIRST) SN[base + k] = shN[k];
    }
    if constexpr (WRITE_RHS) {
      __syncthreads();
      if (live) {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) shS[t * NEQ + v] = rhs[v];
      }
      __syncthreads();
      for (int k = t; k < nval; k += FV_BLOCK)
        if (!(cm && cm[i0 + k / NEQ])) RHS[base + k] = shS[k];
    }
  } else if constexpr (WRITE_RHS) {
    for (int k = t; k < nval; k += FV_BLOCK)
      if (!(cm && cm[i0 + k / NEQ])) RHS[base + k] = shS[k];
  }
}

/* standalone RK(irk) from the RHS array (RK.c:16-51) */
template <int NEQ>
__global__ __launch_bounds__(FV_BLOCK) void k_rk(DevMesh M, double *__restrict__ S,
                                                 double *__restrict__ SN,
                                                 const double *__restrict__ RHS, int irk,
                                                 double dt) {
  const int i = logical_block() * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  const double Vol = M.vol[(M.cflag[i] >> 16) & 0xff];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const size_t o = (size_t)i * NEQ + v;
    const double s = S[o];
    if (irk == 0) SN[o] = s;
    else if (irk >= 1 && irk <= 4) S[o] = rk_update(irk, s, SN[o], RHS[o], dt, Vol);
  }
}

/* per-cell sum_a (|u_a| + c)/h_a, wavefront max by lane shuffles, block max in
 * LDS, one 64-bit integer atomicMax per block on the (non-negative) double bits */
template <int MODEL>
__global__ __launch_bounds__(FV_BLOCK) void k_cfl(DevMesh M, DevParams P,
                                                  const double *__restrict__ S,
                                                  unsigned long long *__restrict__ out,
                                                  int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  __shared__ double red[FV_BLOCK / 64];
  const int i = blockIdx.x * FV_BLOCK + threadIdx.x;
  double m = 0.0;
  if (i < M.N) {
    double U[NEQ];
    load_S<NEQ>(S, M.ns, i, U);
    double r, u[3], e, p, c, y;
    eos64<MODEL>(P, U, false, r, u, e, p, c, y);
    const int L = (M.cflag[i] >> 16) & 0xff;
    const double *h = M.h + 3 * L;
    m = (fabs(u[0]) + c) / h[0] + (fabs(u[1]) + c) / h[1] + (fabs(u[2]) + c) / h[2];
    if (!(m == m) || !(r > 0.0)) { atomicMin(err, i); m = 0.0; }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < FV_BLOCK / 64; ++w) b = fmax(b, red[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(b));
  }
}

/* fixed-order totals: block b sums cells [b*chunk, (b+1)*chunk) thread-strided,
 * tree-reduces in LDS, writes part[b][neq+1]; a second pass sums the parts */
template <int NEQ>
__global__ __launch_bounds__(FV_BLOCK) void k_totals_part(DevMesh M, const double *__restrict__ S,
                                                          double *__restrict__ part, int chunk) {
  __shared__ double sh[NEQ + 1][FV_BLOCK];
  const int b = blockIdx.x, t = threadIdx.x;
  double acc[NEQ + 1];
#pragma unroll
  for (int v = 0; v <= NEQ; ++v) acc[v] = 0.0;
  const int i0 = b * chunk, i1 = min(M.N, i0 + chunk);
  for (int i = i0 + t; i < i1; i += FV_BLOCK) {
    const double V = M.vol[(M.cflag[i] >> 16) & 0xff];
#pragma unroll
    for (int v = 0; v < NEQ; ++v) acc[v] += S[(size_t)i * NEQ + v] * V;
    acc[NEQ] += V;
  }
#pragma unroll
  for (int v = 0; v <= NEQ; ++v) sh[v][t] = acc[v];
  __syncthreads();
  for (int w = FV_BLOCK / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int v = 0; v <= NEQ; ++v) sh[v][t] += sh[v][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int v = 0; v <= NEQ; ++v) part[(size_t)b * (NEQ + 1) + v] = sh[v][0];
  }
}

template <int NEQ>
__global__ void k_totals_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int v = 0; v <= NEQ; ++v) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * (NEQ + 1) + v];
    out[v] = s;
  }
}

template <int NEQ>
__global__ __launch_bounds__(FV_BLOCK) void k_check(DevMesh M, const double *__restrict__ S,
                                                    const double *__restrict__ RHS,
                                                    int *__restrict__ err) {
  const int i = blockIdx.x * FV_BLOCK + threadIdx.x;
  if (i >= M.N) return;
  bool bad = !(S[(size_t)i * NEQ] > 0.0);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    double s = S[(size_t)i * NEQ + v], r = RHS[(size_t)i * NEQ + v];
    bad |= !(s - s == 0.0) || !(r - r == 0.0);
  }
  if (bad) atomicMin(err, i);
}

__global__ void k_fill_int(int *p, int n, int val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long val) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = val;
}

static void to_dev_params(const fvg_params *p, DevParams *d) {
  d->model = p->model;
  d->ns = p->ns;
  d->ischeme = p->ischeme;
  d->limiter = p->limiter;
  d->visc_work = p->visc_work;
  d->neq = fvg_neq(p->model);
  d->gamma = p->gamma;
  d->pinf = p->pinf;
  d->mu = p->mu;
  d->ymin = p->ymin;
  d->baro_nexp = p->baro_nexp;
  d->baro_beta = p->baro_beta;
  d->baro_cbar = p->baro_cbar;
  d->baro_psat = p->baro_psat;
  d->baro_rsat = p->baro_rsat;
  d->mix_RG = p->mix_RG;
  d->mix_Tref = p->mix_Tref;
  d->mix_rsat = p->mix_rsat;
  d->mix_psat = p->mix_psat;
  d->mix_cliq = p->mix_cliq;
  d->mix_cmix = p->mix_cmix;
  for (int v = 0; v < 5; ++v) d->inlet[v] = p->inlet[v];
  d->va_eps = p->vanalbada_eps;
  d->iprimtv = p->iprimtv;
  d->pvrs = p->wave_speeds == FVG_WAVES_PVRS;
  d->kheat = p->prandtl > 0.0 ? p->mu * p->gamma / ((p->gamma - 1.0) * p->prandtl) : 0.0;
}

/* the options beyond the reference's behaviour as OPT_* bits: the kernels' option
 * variants run where any is set */
static int opt_mask(const fvg_params &p) {
  return (p.iprimtv ? OPT_PRIM : 0) | (p.wave_speeds != FVG_WAVES_DAVIS ? OPT_PVRS : 0) |
         (p.prandtl > 0.0 ? OPT_HEAT : 0);
}
static bool has_opts(const fvg_params &p) { return opt_mask(p) != 0; }

static int check_params(const fvg_params *p) {
  if (!p) return FVG_ERR_ARG;
  if (fvg_neq(p->model) < 0) return FVG_ERR_UNSUPPORTED;
  if (p->ischeme != 1 && p->ischeme != 2) return FVG_ERR_UNSUPPORTED;
  if (p->limiter != FVG_MINMOD && p->limiter != FVG_VANALBADA) return FVG_ERR_UNSUPPORTED;
  if (p->precision != FVG_PLAIN && p->precision != FVG_MIXED) return FVG_ERR_UNSUPPORTED;
  if (p->ns && p->model == 2) return FVG_ERR_UNSUPPORTED; /* viscous path needs 5 rows of SG */
  if (p->iprimtv != 0 && p->iprimtv != 1) return FVG_ERR_ARG;
  if (p->wave_speeds != FVG_WAVES_DAVIS && p->wave_speeds != FVG_WAVES_PVRS) return FVG_ERR_ARG;
  /* the PVRS estimate and the heat flux are written for the ideal / stiffened gas */
  if (p->wave_speeds == FVG_WAVES_PVRS && p->model != 1) return FVG_ERR_UNSUPPORTED;
  if (!(p->prandtl >= 0.0)) return FVG_ERR_ARG;
  if (p->prandtl > 0.0 && (p->model != 1 || !p->ns)) return FVG_ERR_UNSUPPORTED;
  return FVG_OK;
}

/* a block size rounded up to an eighth of its leading power of two */
static size_t block_class(size_t b) {
  if (b <= 4096) return 4096;
  const size_t q = (size_t)1 << (63 - __builtin_clzll((unsigned long long)b) - 3);
  return (b + q - 1) / q * q;
}

hipError_t fv_dmalloc_raw(fvg_ctx *c, void **p, size_t bytes, bool slack) {
  *p = nullptr;
  if (!bytes) bytes = 1;
  int best = -1;
  for (int k = 0; k < (int)c->dcache.size(); ++k) {
    const size_t cap = c->dcache[k].second;
    if (cap >= bytes && cap <= 2 * bytes + 4096 && (best < 0 || cap < c->dcache[best].second)) best = k;
  }
  if (best >= 0) {
    c->dlive.push_back(c->dcache[best]);
    c->dcache.erase(c->dcache.begin() + best);
    *p = c->dlive.back().first;
    return hipSuccess;
  }
  const size_t cap = block_class(slack ? bytes + bytes / 4 : bytes);
  hipError_t e = hipMalloc(p, cap);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fv_dcache_clear(c);
    e = hipMalloc(p, cap);
  }
  if (e == hipSuccess) c->dlive.emplace_back(*p, cap);
  else *p = nullptr;
  return e;
}

void fv_dfree(fvg_ctx *c, void *p) {
  if (!p) return;
  for (size_t k = 0; k < c->dlive.size(); ++k)
    if (c->dlive[k].first == p) {
      c->dcache.push_back(c->dlive[k]);
      c->dlive.erase(c->dlive.begin() + (long)k);
      return;
    }
  (void)hipFree(p);
}

void fv_dcache_trim(fvg_ctx *c) {
  size_t live = 0, cached = 0;
  for (auto &b : c->dlive) live += b.second;
  for (auto &b : c->dcache) cached += b.second;
  while (!c->dcache.empty() && cached > live) {
    cached -= c->dcache.front().second;
    (void)hipFree(c->dcache.front().first);
    c->dcache.erase(c->dcache.begin());
  }
}

void fv_dcache_clear(fvg_ctx *c) {
  for (auto &b : c->dcache) (void)hipFree(b.first);
  c->dcache.clear();
}

static void release_step_graph(fvg_ctx *c) {
  if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
  c->step_graph = nullptr;
  c->step_graph_key.clear();
}

static void free_mesh(fvg_ctx *c) {
  fv_adj_release(c);
  fv_tiles_release(c);
  release_step_graph(c);
  void *ptrs[] = {c->cflag, c->nbr, c->quad, c->fL, c->fR, c->fLL, c->fRR, c->gslot, c->fmeta, c->tab,
                  c->S, c->SN, c->Sx, c->RHS, c->SG, c->SF, c->SGm, c->G, c->Sf, c->Sfx, c->keys,
                  c->hkeys, c->hval};
  for (void *p : ptrs) fv_dfree(c, p);
  c->cflag = nullptr; c->nbr = c->quad = c->fL = c->fR = c->fLL = c->fRR = c->gslot = nullptr;
  c->fmeta = nullptr; c->tab = nullptr;
  c->S = c->SN = c->RHS = c->SG = nullptr;
  c->SF = c->SGm = c->G = nullptr;
  c->Sf = c->Sfx = nullptr;
  c->Sx = nullptr;
  c->stage_in = nullptr;
  c->keys = c->hkeys = nullptr;
  c->hval = nullptr;
  c->hbits = 0;
  c->sf_valid = false;
  c->has_mesh = false;
}

/* Frees the previous mesh and allocates the arrays of a mesh of N cells, F faces, Q
 * quads and nl table rows; the state arrays are zeroed (padding included) so that no
 * kernel reads uninitialised memory.  topo = false leaves cflag / nbr to the caller
 * (the device re-flatten hands over the ones it built). */
int fv_alloc_mesh(fvg_ctx *c, int N, int F, int Q, int nl, const int32_t nfa[3],
                  const int32_t nreg[3], bool topo) {
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  c->N = N;
  c->ns = (int)round_up(N, 64);
  c->F = F;
  c->nfp = (int)round_up(F, 64);
  c->Q = Q;
  c->nlevel = nl;
  for (int a = 0; a < 3; ++a) {
    c->nfa[a] = nfa[a];
    c->nreg[a] = nreg[a];
  }
  const int ns = c->ns, neq = c->neq;
  const size_t rs = c->rsize;
  const bool slack = !topo; /* a mesh rebuilt by adaptation may grow next time */
  if (topo) {
    HIPCHK(c, fv_dmalloc(c, &c->cflag, (size_t)ns * 4, slack));
    HIPCHK(c, fv_dmalloc(c, &c->nbr, (size_t)6 * ns * 4, slack));
  }
  HIPCHK(c, fv_dmalloc(c, &c->gslot, (size_t)6 * ns * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->quad, (size_t)4 * (Q ? Q : 1) * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->fL, (size_t)c->nfp * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->fR, (size_t)c->nfp * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->fLL, (size_t)c->nfp * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->fRR, (size_t)c->nfp * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->fmeta, (size_t)c->nfp * 4, slack));
  HIPCHK(c, fv_dmalloc(c, &c->tab, (size_t)34 * nl * 8, slack));
  HIPCHK(c, fv_dmalloc(c, &c->S, (size_t)neq * ns * 8, slack));
  HIPCHK(c, fv_dmalloc(c, &c->SN, (size_t)neq * ns * 8, slack));
  HIPCHK(c, fv_dmalloc(c, &c->RHS, (size_t)neq * ns * 8, slack));
  HIPCHK(c, fv_dmalloc(c, &c->SG, (size_t)(neq + 1) * 3 * ns * 8, slack));
  HIPCHK(c, fv_dmalloc(c, &c->SF, (size_t)6 * neq * ns * rs, slack));
  /* 9 momentum gradients, then 3 of p / rho (heat conduction option) */
  HIPCHK(c, fv_dmalloc(c, &c->SGm, (size_t)VROWS * ns * rs, slack));
  HIPCHK(c, fv_dmalloc(c, &c->G, (size_t)neq * c->nfp * rs, slack));
  HIPCHK(c, fv_dmalloc(c, &c->Sx, (size_t)neq * ns * 8, slack));
  if (rs == 4 && c->prm.model == 1) {
    HIPCHK(c, fv_dmalloc(c, &c->Sf, (size_t)neq * ns * 4, slack));
    HIPCHK(c, fv_dmalloc(c, &c->Sfx, (size_t)neq * ns * 4, slack));
  }
  c->sf_valid = false;
  c->stage_in = nullptr;
  hipStream_t st = c->stream;
  HIPCHK(c, hipMemsetAsync(c->S, 0, (size_t)neq * ns * 8, st));
  HIPCHK(c, hipMemsetAsync(c->SN, 0, (size_t)neq * ns * 8, st));
  HIPCHK(c, hipMemsetAsync(c->Sx, 0, (size_t)neq * ns * 8, st));
  HIPCHK(c, hipMemsetAsync(c->RHS, 0, (size_t)neq * ns * 8, st));
  HIPCHK(c, hipMemsetAsync(c->SG, 0, (size_t)(neq + 1) * 3 * ns * 8, st));
  HIPCHK(c, hipMemsetAsync(c->SF, 0, (size_t)6 * neq * ns * rs, st));
  HIPCHK(c, hipMemsetAsync(c->SGm, 0, (size_t)VROWS * ns * rs, st));
  HIPCHK(c, hipMemsetAsync(c->G, 0, (size_t)neq * c->nfp * rs, st));
  fv_dcache_trim(c);
  return FVG_OK;
}

/* Per-level tables (h, distance() factors, areas, volumes and the reciprocals of
 * the d_cc factors) into c->tab, nl rows each. */
int fv_upload_tables(fvg_ctx *c, int nl, const double *h, const double *dist, const double *area,
                     const double *vol) {
  std::vector<double> tab((size_t)34 * nl, 0.0);
  memcpy(tab.data(), h, (size_t)3 * nl * 8);
  memcpy(tab.data() + 3 * nl, dist, (size_t)18 * nl * 8);
  memcpy(tab.data() + 21 * nl, area, (size_t)3 * nl * 8);
  memcpy(tab.data() + 24 * nl, vol, (size_t)nl * 8);
  /* reciprocals of the d_cc factors (a zero distance gives inf, as 1/0 does) */
  for (int k = 0; k < 9 * nl; ++k) {
    const double dcc = dist[2 * k + 1];
    tab[25 * nl + k] = dcc != 0.0 ? 1.0 / dcc : HUGE_VAL;
  }
  HIPCHK(c, hipMemcpyAsync(c->tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream));
  /* the host vector must outlive the copy */
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

extern "C" {

/* Staging of the refinement flags of fvg_adapt_flags_static: a file-scope device array,
 * one copy on every GPU the library runs on, held until the process ends, so that the
 * entry point allocates nothing.  Host code never passes the symbol itself to a kernel
 * or a copy (on the host it is a stand-in, not a device address): it uses the address
 * hipGetSymbolAddress returns for the current device, looked up once per device. */
static __device__ __align__(256) int8_t scratch[FVG_STATIC_FLAGS_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices];
int8_t *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
}  // namespace

/* the device address of `scratch` on `device`, which must be current */
static hipError_t scratch_address(int device, int8_t **ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void *a = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? (int8_t *)a : nullptr;
    g_scratch_done[device] = true;
  }
  *ptr = g_scratch_ptr[device];
  return g_scratch_err[device];
}

int fvg_create(int device, const fvg_params *p, fvg_ctx **out) {
  if (!out) return FVG_ERR_ARG;
  int rc = check_params(p);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FVG_ERR_HIP;
  if (device < 0 || device >= ndev) return FVG_ERR_ARG;
  fvg_ctx *c = new fvg_ctx();
  c->device = device;
  c->prm = *p;
  c->neq = fvg_neq(p->model);
  c->rsize = p->precision == FVG_MIXED ? 4 : 8;
  to_dev_params(p, &c->dp);
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->err, 4 * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->redu, 64) != hipSuccess) {
    delete c;
    return FVG_ERR_HIP;
  }
  /* with the device current; a failure is reported by fvg_adapt_flags_static only */
  c->flag_scratch_rc = (int)scratch_address(device, &c->flag_scratch);
  *out = c;
  return FVG_OK;
}

int fvg_destroy(fvg_ctx *c) {
  if (!c) return FVG_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  fv_remesh_release(c);
  free_mesh(c);
  fv_dcache_clear(c);
  if (c->err) (void)hipFree(c->err);
  if (c->red) (void)hipFree(c->red);
  if (c->redu) (void)hipFree(c->redu);
  if (c->ev_ok)
    for (auto &e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FVG_OK;
}

int fvg_set_params(fvg_ctx *c, const fvg_params *p) {
  if (!c) return FVG_ERR_ARG;
  int rc = check_params(p);
  if (rc) return set_err(c, rc, "unsupported parameters");
  if (fvg_neq(p->model) != c->neq) return set_err(c, FVG_ERR_ARG, "model changes neq");
  size_t rs = p->precision == FVG_MIXED ? 4 : 8;
  if (c->has_mesh && rs != c->rsize) return set_err(c, FVG_ERR_STATE, "precision fixed by set_mesh");
  c->prm = *p;
  c->rsize = rs;
  to_dev_params(p, &c->dp);
  return FVG_OK;
}

int fvg_ncell(const fvg_ctx *c) { return c ? c->N : FVG_ERR_ARG; }
int fvg_set_tiling(fvg_ctx *c, int mode) {
  if (!c) return FVG_ERR_ARG;
  if (mode != FVG_TILING_OFF && mode != FVG_TILING_ON && mode != FVG_TILING_AUTO)
    return set_err(c, FVG_ERR_ARG, "tiling mode must be FVG_TILING_OFF, _ON or _AUTO");
  c->tiling = mode;
  return FVG_OK;
}
int fvg_tile_count(const fvg_ctx *c) { return c ? c->ntile : FVG_ERR_ARG; }
int fvg_set_graphs(fvg_ctx *c, int on) {
  if (!c) return FVG_ERR_ARG;
  c->graphs = on != 0;
  if (!c->graphs) release_step_graph(c);
  return FVG_OK;
}
void *fvg_stream(fvg_ctx *c) { return c ? (void *)c->stream : nullptr; }
const char *fvg_last_error(const fvg_ctx *c) { return c ? c->last.c_str() : "null context"; }

int fvg_set_mesh(fvg_ctx *c, const fvg_mesh *m) {
  if (!c || !m) return FVG_ERR_ARG;
  if (m->neq != c->neq) return set_err(c, FVG_ERR_ARG, "mesh neq differs from the model's");
  if (m->ncell <= 0 || m->nface <= 0 || m->nlevel <= 0 || m->nlevel > 32 || m->nquad < 0)
    return set_err(c, FVG_ERR_ARG, "bad mesh sizes");
  if ((int64_t)m->nface_axis[0] + m->nface_axis[1] + m->nface_axis[2] != m->nface)
    return set_err(c, FVG_ERR_ARG, "face axis counts do not sum to nface");
  const int N = m->ncell, F = m->nface, Q = m->nquad, nl = m->nlevel;
  /* host-side validation of every index the kernels dereference */
  for (int i = 0; i < N; ++i) {
    uint32_t cf = m->cflag[i];
    int L = (cf >> 16) & 0xff;
    if (L < 1 || L + 1 >= nl + 1) return set_err(c, FVG_ERR_ARG, "cell level outside tables");
    for (int f = 0; f < 6; ++f) {
      int cls = (cf >> (2 * f)) & 3;
      int nb = m->nbr[(size_t)f * N + i];
      if (cls == CLS_SAME || cls == CLS_COARSER) {
        if (nb < 0 || nb >= N) return set_err(c, FVG_ERR_ARG, "neighbour index out of range");
      } else if (cls == CLS_FINER) {
        if (nb < 0 || nb >= Q) return set_err(c, FVG_ERR_ARG, "quad index out of range");
        if (L + 1 >= nl) return set_err(c, FVG_ERR_ARG, "finer level outside tables");
      } else {
        if (nb != -2 && nb != -3) return set_err(c, FVG_ERR_ARG, "boundary type not 1 or 2");
      }
      int g = m->gslot[(size_t)f * N + i];
      if (g < -1 || (g >= 0 && (g >> 1) + (cls == CLS_FINER ? 3 : 0) >= F))
        return set_err(c, FVG_ERR_ARG, "gather slot out of range");
    }
  }
  for (int64_t q = 0; q < 4LL * Q; ++q)
    if (m->quad_nbr[q] < 0 || m->quad_nbr[q] >= N) return set_err(c, FVG_ERR_ARG, "bad quad neighbour");
  /* regular faces: each axis group starts with nface_reg[a] faces whose stencil
   * face_LL, face_L, face_R, face_RR is four same-level cells along the axis */
  {
    static const int opp[6] = {2, 3, 0, 1, 5, 4};
    int64_t start = 0;
    for (int a = 0; a < 3; ++a) {
      if (m->nface_reg[a] < 0 || m->nface_reg[a] > m->nface_axis[a])
        return set_err(c, FVG_ERR_ARG, "bad regular face count");
      if (m->nface_reg[a] && (!m->face_LL || !m->face_RR))
        return set_err(c, FVG_ERR_ARG, "regular faces without face_LL / face_RR");
      for (int64_t q = start; q < start + m->nface_reg[a]; ++q) {
        const int iL = m->face_L[q], iR = m->face_R[q], iLL = m->face_LL[q], iRR = m->face_RR[q];
        const uint32_t mt = m->face_meta[q];
        const int f = mt & 7, fo = (mt >> 3) & 7;
        if (iL < 0 || iL >= N || iR < 0 || iR >= N || iLL < 0 || iLL >= N || iRR < 0 || iRR >= N ||
            f > 5 || fo > 5 || ((mt >> 6) & 3) != 0)
          return set_err(c, FVG_ERR_ARG, "bad regular face");
        const uint32_t cL = m->cflag[iL], cR = m->cflag[iR];
        if (((cL >> (2 * f)) & 3) != CLS_SAME || ((cL >> (2 * opp[f])) & 3) != CLS_SAME ||
            ((cR >> (2 * opp[fo])) & 3) != CLS_SAME || m->nbr[(size_t)f * N + iL] != iR ||
            m->nbr[(size_t)opp[f] * N + iL] != iLL || m->nbr[(size_t)opp[fo] * N + iR] != iRR ||
            c_face_axis_host(f) != a)
          return set_err(c, FVG_ERR_ARG, "face listed as regular does not have a regular stencil");
      }
      start += m->nface_axis[a];
    }
  }
  for (int j = 0; j < F; ++j) {
    if (m->face_L[j] < 0 || m->face_L[j] >= N || m->face_R[j] < 0 || m->face_R[j] >= N)
      return set_err(c, FVG_ERR_ARG, "face cell out of range");
    uint32_t mt = m->face_meta[j];
    int lev = (mt >> 8) & 0xff;
    if ((mt & 7) > 5 || ((mt >> 3) & 7) > 5 || lev < 1 || lev >= nl || ((mt >> 6) & 3) == 3)
      return set_err(c, FVG_ERR_ARG, "bad face meta");
  }
  int rc = fv_alloc_mesh(c, N, F, Q, nl, m->nface_axis, m->nface_reg, true);
  if (rc) return rc;
  const int ns = c->ns;
  hipStream_t st = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->cflag, m->cflag, (size_t)N * 4, hipMemcpyHostToDevice, st));
  for (int f = 0; f < 6; ++f) {
    HIPCHK(c, hipMemcpyAsync(c->nbr + (size_t)f * ns, m->nbr + (size_t)f * N, (size_t)N * 4,
                             hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->gslot + (size_t)f * ns, m->gslot + (size_t)f * N, (size_t)N * 4,
                             hipMemcpyHostToDevice, st));
  }
  if (Q) HIPCHK(c, hipMemcpyAsync(c->quad, m->quad_nbr, (size_t)4 * Q * 4, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->fL, m->face_L, (size_t)F * 4, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->fR, m->face_R, (size_t)F * 4, hipMemcpyHostToDevice, st));
  if (m->face_LL && m->face_RR) {
    HIPCHK(c, hipMemcpyAsync(c->fLL, m->face_LL, (size_t)F * 4, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->fRR, m->face_RR, (size_t)F * 4, hipMemcpyHostToDevice, st));
  } else {
    HIPCHK(c, hipMemsetAsync(c->fLL, 0xff, (size_t)F * 4, st));
    HIPCHK(c, hipMemsetAsync(c->fRR, 0xff, (size_t)F * 4, st));
  }
  HIPCHK(c, hipMemcpyAsync(c->fmeta, m->face_meta, (size_t)F * 4, hipMemcpyHostToDevice, st));
  rc = fv_upload_tables(c, nl, m->h, m->dist, m->area, m->vol);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(st));
  rc = fv_tiles_build(c, m);
  if (rc) return rc;
  c->has_mesh = true;
  return FVG_OK;
}

}  // extern "C"

/* ------------------------------------------------------------------ field copies */

static size_t field_rows(const fvg_ctx *c, int which, int *len, void **dptr, bool *isR) {
  *isR = false;
  switch (which) {
    case FVG_FIELD_S: *dptr = c->S; *len = c->ns; return c->neq;
    case FVG_FIELD_SN: *dptr = c->SN; *len = c->ns; return c->neq;
    case FVG_FIELD_RHS: *dptr = c->RHS; *len = c->ns; return c->neq;
    case FVG_FIELD_SF: *dptr = c->SF; *len = c->ns; *isR = true; return 6 * c->neq;
    case FVG_FIELD_SG: *dptr = c->SG; *len = c->ns; return 3 * (c->neq + 1);
    case FVG_FIELD_FACEFLUX: *dptr = c->G; *len = c->nfp; *isR = true; return c->neq;
    case FVG_FIELD_STAGE_IN: *dptr = c->stage_in; *len = c->ns; return c->neq;
    default: return 0;
  }
}

/* S, SN, RHS and the face flux are stored cell- (face-) major on the device */
static bool cell_major(int which) {
  return which == FVG_FIELD_S || which == FVG_FIELD_SN || which == FVG_FIELD_RHS ||
         which == FVG_FIELD_FACEFLUX || which == FVG_FIELD_STAGE_IN;
}

static int copy_field(fvg_ctx *c, int which, double *host, bool to_dev) {
  if (!c || !host) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  int len;
  void *d;
  bool isR;
  size_t rows = field_rows(c, which, &len, &d, &isR);
  if (!rows) return set_err(c, FVG_ERR_ARG, "unknown field");
  if (which == FVG_FIELD_STAGE_IN) {
    if (to_dev) return set_err(c, FVG_ERR_ARG, "the last stage's input is read only");
    if (!d) return set_err(c, FVG_ERR_STATE, "no step has run since the state was last written");
  }
  const int n = which == FVG_FIELD_FACEFLUX ? c->F : c->N;
  (void)hipSetDevice(c->device);
  if (to_dev && which == FVG_FIELD_S) c->sf_valid = false;
  if (to_dev) c->stage_in = nullptr;
  const size_t es = isR ? c->rsize : 8;
  if (cell_major(which)) {
    /* host var-major [rows][n] <-> device [n][rows] */
    std::vector<char> tmp(rows * (size_t)n * es);
    if (to_dev) {
      for (size_t r = 0; r < rows; ++r)
        for (int i = 0; i < n; ++i) {
          const double x = host[r * (size_t)n + i];
          if (es == 8) ((double *)tmp.data())[(size_t)i * rows + r] = x;
          else ((float *)tmp.data())[(size_t)i * rows + r] = (float)x;
        }
      HIPCHK(c, hipMemcpyAsync(d, tmp.data(), tmp.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
      HIPCHK(c, hipMemcpyAsync(tmp.data(), d, tmp.size(), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t r = 0; r < rows; ++r)
        for (int i = 0; i < n; ++i)
          host[r * (size_t)n + i] = es == 8 ? ((double *)tmp.data())[(size_t)i * rows + r]
                                            : (double)((float *)tmp.data())[(size_t)i * rows + r];
    }
    return FVG_OK;
  }
  std::vector<char> tmp((size_t)n * es);
  for (size_t r = 0; r < rows; ++r) {
    char *dp = (char *)d + r * (size_t)len * es;
    double *hp = host + r * (size_t)n;
    if (to_dev) {
      if (es == 8) memcpy(tmp.data(), hp, (size_t)n * 8);
      else for (int i = 0; i < n; ++i) ((float *)tmp.data())[i] = (float)hp[i];
      HIPCHK(c, hipMemcpyAsync(dp, tmp.data(), (size_t)n * es, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
      HIPCHK(c, hipMemcpyAsync(tmp.data(), dp, (size_t)n * es, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (es == 8) memcpy(hp, tmp.data(), (size_t)n * 8);
      else for (int i = 0; i < n; ++i) hp[i] = (double)((float *)tmp.data())[i];
    }
  }
  return FVG_OK;
}

/* var-major [neq][N] <-> cell-major [N][neq] on the device */
__global__ void k_transpose_state(const double *__restrict__ src, double *__restrict__ dst, int N,
                                  int neq, int to_cell_major) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int v = 0; v < neq; ++v) {
    if (to_cell_major) dst[(size_t)i * neq + v] = src[(size_t)v * N + i];
    else dst[(size_t)v * N + i] = src[(size_t)i * neq + v];
  }
}

/* ------------------------------------------------------------------ launches */

static inline int nblocks(int n) { return (n + FV_BLOCK - 1) / FV_BLOCK; }
/* The gather's grid, padded to a multiple of the 8 XCDs so that logical_block() hands each
 * XCD one contiguous range of cells (with any other count it keeps the hardware's
 * round-robin order, and the face fluxes shared by neighbouring blocks are fetched into
 * several XCDs' L2). */
static inline int gather_grid(int n) { return (int)round_up(nblocks(n), 8); }

/* Every launcher takes the state it reads (S) explicitly: c->S in the time loop,
 * a segment slot when the adjoint recomputes stage states. */
template <int MODEL>
static void launch_recon(fvg_ctx *c, const double *S) {
  DevMesh M = dev_mesh(c);
  dim3 g(nblocks(c->N)), b(FV_BLOCK);
  const bool prim = c->prm.iprimtv != 0;
  if (c->rsize == 8) {
    if (prim)
      hipLaunchKernelGGL((k_recon<MODEL, double, true>), g, b, 0, c->stream, M, c->dp, S, (double *)c->SF);
    else
      hipLaunchKernelGGL((k_recon<MODEL, double>), g, b, 0, c->stream, M, c->dp, S, (double *)c->SF);
  } else {
    if (prim)
      hipLaunchKernelGGL((k_recon<MODEL, float, true>), g, b, 0, c->stream, M, c->dp, S, (float *)c->SF);
    else
      hipLaunchKernelGGL((k_recon<MODEL, float>), g, b, 0, c->stream, M, c->dp, S, (float *)c->SF);
  }
}

template <int MODEL>
static void launch_grad(fvg_ctx *c, const double *S, bool full, bool raw = false,
                        const float *Sf = nullptr) {
  DevMesh M = dev_mesh(c);
  dim3 g(nblocks(c->N)), b(FV_BLOCK);
  const uint8_t *skip = full ? nullptr : c->stage_gskip, *cmask = full ? nullptr : c->stage_cmask;
  if (Sf && !full && !raw && c->rsize == 4) {
    /* the mixed ideal-gas stage: the viscous rows from the FP32 copy of S */
    if constexpr (MODEL == 1) {
      if (c->dp.kheat != 0.0)
        hipLaunchKernelGGL((k_grad<MODEL, false, float, true, false, float>), g, b, 0, c->stream, M,
                           c->dp, Sf, c->SG, (float *)c->SGm, skip, cmask);
      else
        hipLaunchKernelGGL((k_grad<MODEL, false, float, false, false, float>), g, b, 0, c->stream, M,
                           c->dp, Sf, c->SG, (float *)c->SGm, skip, cmask);
      return;
    }
  }
  if (full)
    hipLaunchKernelGGL((k_grad<MODEL, true, double>), g, b, 0, c->stream, M, c->dp, S, c->SG,
                       (double *)nullptr, skip, cmask);
  else if (raw) {
    if (c->rsize == 8)
      hipLaunchKernelGGL((k_grad<MODEL, false, double, false, true>), g, b, 0, c->stream, M, c->dp,
                         S, c->SG, (double *)c->SGm, skip, cmask);
    else
      hipLaunchKernelGGL((k_grad<MODEL, false, float, false, true>), g, b, 0, c->stream, M, c->dp,
                         S, c->SG, (float *)c->SGm, skip, cmask);
  } else if (MODEL == 1 && c->dp.kheat != 0.0) {
    if constexpr (MODEL == 1) {
      if (c->rsize == 8)
        hipLaunchKernelGGL((k_grad<MODEL, false, double, true>), g, b, 0, c->stream, M, c->dp, S,
                           c->SG, (double *)c->SGm, skip, cmask);
      else
        hipLaunchKernelGGL((k_grad<MODEL, false, float, true>), g, b, 0, c->stream, M, c->dp, S,
                           c->SG, (float *)c->SGm, skip, cmask);
    }
  } else if (c->rsize == 8)
    hipLaunchKernelGGL((k_grad<MODEL, false, double>), g, b, 0, c->stream, M, c->dp, S, c->SG,
                       (double *)c->SGm, skip, cmask);
  else
    hipLaunchKernelGGL((k_grad<MODEL, false, float>), g, b, 0, c->stream, M, c->dp, S, c->SG,
                       (float *)c->SGm, skip, cmask);
}

/* The option variant a parameter set runs: none, the heat-only one (the viscous
 * Taylor-Green configuration: no other option code compiled in), or the one with all
 * option bits (their runtime switches pick the path). */
template <int MODEL, bool NS>
static int opt_variant(const fvg_ctx *c) {
  const int m = opt_mask(c->prm);
  if (m == 0) return 0;
  if (m == OPT_HEAT && MODEL == 1 && NS) return OPT_HEAT;
  return OPT_ALL;
}

template <int MODEL, bool NS, typename R, int OPT>
static void launch_face_t(fvg_ctx *c, const double *S) {
  DevMesh M = dev_mesh(c);
  const int n0 = c->nfa[0], n1 = c->nfa[1], n2 = c->nfa[2];
  const int b0 = nblocks(n0), b1 = nblocks(n1), b2 = nblocks(n2);
  dim3 g(b0 + b1 + b2), b(FV_BLOCK);
  hipLaunchKernelGGL((k_face<MODEL, NS, R, R, OPT>), g, b, 0, c->stream, M, c->dp, b0, b1, n0, n1,
                     n2, S, (const R *)c->SF, (const R *)c->SGm, (R *)c->G, c->err);
}

template <int MODEL, bool NS, typename R>
static void launch_face_r(fvg_ctx *c, const double *S) {
  const int v = opt_variant<MODEL, NS>(c);
  if (v == 0) launch_face_t<MODEL, NS, R, 0>(c, S);
  else if (v == OPT_HEAT) launch_face_t<MODEL, NS, R, OPT_HEAT>(c, S);
  else launch_face_t<MODEL, NS, R, OPT_ALL>(c, S);
}

template <int MODEL, bool NS>
static void launch_face_m(fvg_ctx *c, const double *S) {
  if (c->rsize == 8) launch_face_r<MODEL, NS, double>(c, S);
  else launch_face_r<MODEL, NS, float>(c, S);
}

template <int MODEL, bool NS, typename R, int OPT>
static void launch_face_fused_t(fvg_ctx *c, const double *S) {
  DevMesh M = dev_mesh(c);
  const int n0 = c->nfa[0], n1 = c->nfa[1], n2 = c->nfa[2];
  const int b0 = nblocks(n0), b1 = nblocks(n1), b2 = nblocks(n2);
  dim3 g(b0 + b1 + b2), b(FV_BLOCK);
  hipLaunchKernelGGL((k_face_fused<MODEL, NS, R, R, OPT>), g, b, 0, c->stream, M, c->dp, n0, n1, n2,
                     c->nreg[0], c->nreg[1], c->nreg[2], S, (const R *)c->SGm, (R *)c->G, c->err,
                     c->stage_fskip);
}

template <int MODEL, bool NS, typename R>
static void launch_face_fused_r(fvg_ctx *c, const double *S) {
  const int v = opt_variant<MODEL, NS>(c);
  if (v == 0) launch_face_fused_t<MODEL, NS, R, 0>(c, S);
  else if (v == OPT_HEAT) launch_face_fused_t<MODEL, NS, R, OPT_HEAT>(c, S);
  else launch_face_fused_t<MODEL, NS, R, OPT_ALL>(c, S);
}

template <int MODEL, bool NS>
static void launch_face_fused_m(fvg_ctx *c, const double *S) {
  if (c->rsize == 8) launch_face_fused_r<MODEL, NS, double>(c, S);
  else launch_face_fused_r<MODEL, NS, float>(c, S);
}

template <int MODEL>
static void launch_face_fused(fvg_ctx *c, const double *S) {
  if (c->prm.ns) launch_face_fused_m<MODEL, true>(c, S);
  else launch_face_fused_m<MODEL, false>(c, S);
}

template <int MODEL>
static void launch_face(fvg_ctx *c, const double *S) {
  if (c->prm.ns) launch_face_m<MODEL, true>(c, S);
  else launch_face_m<MODEL, false>(c, S);
}

/* stage state pointers: the gather reads S (and SN), writes Sout */
struct StagePtrs {
  const double *S;
  double *SN, *Sout;
  float *Sf; /* FP32 copy of Sout to refresh, or null */
};

template <int NEQ, typename R>
static void launch_gather_t(fvg_ctx *c, StagePtrs sp, bool write_rhs, bool do_rk, bool first,
                            int irk, double dt) {
  DevMesh M = dev_mesh(c);
  dim3 g(gather_grid(c->N)), b(FV_BLOCK);
  const R *G = (const R *)c->G;
  if (write_rhs && !do_rk)
    hipLaunchKernelGGL((k_gather_rk<NEQ, R, true, false, false>), g, b, 0, c->stream, M, G, sp.S,
                       sp.SN, c->RHS, irk, dt, c->err, sp.Sout, sp.Sf, c->stage_gskip,
                       c->stage_cmask);
  else if (write_rhs && first)
    hipLaunchKernelGGL((k_gather_rk<NEQ, R, true, true, true>), g, b, 0, c->stream, M, G, sp.S,
                       sp.SN, c->RHS, irk, dt, c->err, sp.Sout, sp.Sf, c->stage_gskip,
                       c->stage_cmask);
  else if (write_rhs)
    hipLaunchKernelGGL((k_gather_rk<NEQ, R, true, true, false>), g, b, 0, c->stream, M, G, sp.S,
                       sp.SN, c->RHS, irk, dt, c->err, sp.Sout, sp.Sf, c->stage_gskip,
                       c->stage_cmask);
  else if (first)
    hipLaunchKernelGGL((k_gather_rk<NEQ, R, false, true, true>), g, b, 0, c->stream, M, G, sp.S,
                       sp.SN, c->RHS, irk, dt, c->err, sp.Sout, sp.Sf, c->stage_gskip,
                       c->stage_cmask);
  else
    hipLaunchKernelGGL((k_gather_rk<NEQ, R, false, true, false>), g, b, 0, c->stream, M, G, sp.S,
                       sp.SN, c->RHS, irk, dt, c->err, sp.Sout, sp.Sf, c->stage_gskip,
                       c->stage_cmask);
}

static void launch_gather(fvg_ctx *c, StagePtrs sp, bool write_rhs, bool do_rk, bool first, int irk,
                          double dt) {
  if (c->neq == 5) {
    if (c->rsize == 8) launch_gather_t<5, double>(c, sp, write_rhs, do_rk, first, irk, dt);
    else launch_gather_t<5, float>(c, sp, write_rhs, do_rk, first, irk, dt);
  } else {
    if (c->rsize == 8) launch_gather_t<4, double>(c, sp, write_rhs, do_rk, first, irk, dt);
    else launch_gather_t<4, float>(c, sp, write_rhs, do_rk, first, irk, dt);
  }
}

static void launch_gather(fvg_ctx *c, bool write_rhs, bool do_rk, bool first, int irk, double dt) {
  launch_gather(c, StagePtrs{c->S, c->SN, c->S, nullptr}, write_rhs, do_rk, first, irk, dt);
}

static void dispatch_recon(fvg_ctx *c, const double *S) {
  switch (c->prm.model) {
    case 1: launch_recon<1>(c, S); break;
    case 2: launch_recon<2>(c, S); break;
    default: launch_recon<3>(c, S); break;
  }
}
static void dispatch_grad(fvg_ctx *c, const double *S, bool full, bool raw = false,
                          const float *Sf = nullptr) {
  switch (c->prm.model) {
    case 1: launch_grad<1>(c, S, full, raw, Sf); break;
    case 2: launch_grad<2>(c, S, full, raw); break;
    default: launch_grad<3>(c, S, full, raw); break;
  }
}
static void dispatch_face_fused(fvg_ctx *c, const double *S) {
  switch (c->prm.model) {
    case 1: launch_face_fused<1>(c, S); break;
    case 2: launch_face_fused<2>(c, S); break;
    default: launch_face_fused<3>(c, S); break;
  }
}
static void dispatch_face(fvg_ctx *c, const double *S) {
  switch (c->prm.model) {
    case 1: launch_face<1>(c, S); break;
    case 2: launch_face<2>(c, S); break;
    default: launch_face<3>(c, S); break;
  }
}

static int reset_err(fvg_ctx *c) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->err, 4, 0x7fffffff);
  return FVG_OK;
}

static int read_err(fvg_ctx *c, int32_t *bad) {
  int h[4];
  HIPCHK(c, hipMemcpyAsync(h, c->err, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *bad = h[0] == 0x7fffffff ? -1 : h[0];
  return FVG_OK;
}

static int finish(fvg_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char b[256];
    snprintf(b, sizeof b, "%s: launch failed: %s", what, hipGetErrorString(e));
    return set_err(c, FVG_ERR_HIP, b);
  }
  return FVG_OK;
}

/* S: the stage's FP64 input (c->Sf holds its FP32 copy) */
static void launch_face_fast(fvg_ctx *c, const double *S) {
  DevMesh M = dev_mesh(c);
  const int n0 = c->nfa[0], n1 = c->nfa[1], n2 = c->nfa[2];
  const int b0 = nblocks(n0), b1 = nblocks(n1), b2 = nblocks(n2);
  dim3 g(b0 + b1 + b2), b(FV_BLOCK);
  if (c->prm.ns && c->dp.kheat != 0.0)
    hipLaunchKernelGGL((k_face_fast<true, true>), g, b, 0, c->stream, M, c->dp, n0, n1, n2,
                       c->nreg[0], c->nreg[1], c->nreg[2], S, c->Sf,
                       (const float *)c->SGm, (float *)c->G, c->err, c->stage_fskip);
  else if (c->prm.ns)
    hipLaunchKernelGGL(k_face_fast<true>, g, b, 0, c->stream, M, c->dp, n0, n1, n2,
                       c->nreg[0], c->nreg[1], c->nreg[2], S, c->Sf,
                       (const float *)c->SGm, (float *)c->G, c->err, c->stage_fskip);
  else
    hipLaunchKernelGGL(k_face_fast<false>, g, b, 0, c->stream, M, c->dp, n0, n1, n2,
                       c->nreg[0], c->nreg[1], c->nreg[2], S, c->Sf,
                       (const float *)c->SGm, (float *)c->G, c->err, c->stage_fskip);
}

/* the FP32 state copy of the mixed ideal-gas path, refreshed if anything else
 * wrote S since the last stage of that path */
static void ensure_sf(fvg_ctx *c) {
  if (c->sf_valid) return;
  const int n = c->N * c->neq;
  hipLaunchKernelGGL(k_to_f32, dim3((n + FV_BLOCK - 1) / FV_BLOCK), dim3(FV_BLOCK), 0, c->stream,
                     (const double *)c->S, c->Sf, n);
  c->sf_valid = true;
}

static const int *stage_list(int method, int *n) {
  static const int st3[3] = {1, 2, 3}, st2[2] = {1, 4}, st1[1] = {1};
  if (method == 3) { *n = 3; return st3; }
  if (method == 2) { *n = 2; return st2; }
  *n = 1;
  return st1;
}

static bool fast_mixed(const fvg_ctx *c) {
  /* heat conduction has its own variant of the FP32-read kernel; the other options run
   * the FP64-read kernels' option variant instead */
  const int m = opt_mask(c->prm);
  return c->Sf && c->prm.model == 1 && c->prm.ischeme == 2 && (m == 0 || m == OPT_HEAT);
}

/* one stage's kernels on the state sp.S -> sp.Sout; ev (optional) receives
 * per-kernel events and kid their kernel ids.  `fast` selects the mixed
 * ideal-gas path (FP32 reconstruction from Sf, which the caller keeps equal to
 * the FP32 copy of sp.S); it needs an out-of-place stage. */
static void stage(fvg_ctx *c, StagePtrs sp, int irk, double dt, bool first, bool write_rhs,
                  hipEvent_t *ev, int *nev, int *kid, bool fast) {
  auto mark = [&](int k) {
    if (ev) { (void)hipEventRecord(ev[*nev], c->stream); kid[*nev] = k; ++*nev; }
  };
  fast = fast && fast_mixed(c) && !first && sp.Sout != sp.S;
  /* Interior tiles: one LDS-tiled kernel updates their cells and the face kernel and the
   * gather skip the blocks that hold only their faces / cells.  Out-of-place stages
   * only: a tile reads its neighbour tiles' S while other workgroups write Sout. */
  const bool tiles = fv_tiles_usable(c) && !first && sp.Sout != sp.S && sp.Sout != sp.SN;
  c->stage_fskip = tiles ? c->face_skip : nullptr;
  c->stage_gskip = tiles ? c->gather_skip : nullptr;
  c->stage_cmask = tiles ? c->tile_cell : nullptr;
  if (c->prm.ns) {
    mark(1);
    /* the mixed ideal-gas form reads its FP32 copy of S for the viscous rows */
    const float *sfin = fast ? c->Sf : nullptr;
    if (!(tiles && c->gather_all_skip)) dispatch_grad(c, sp.S, false, false, sfin);
    if (tiles) { mark(5); fv_launch_tile_grad(c, sp.S, sfin); }
  }
  mark(2);
  /* MUSCL: the reconstruction runs inside the face kernel (no SF round trip) */
  if (!(tiles && c->face_all_skip)) {
    if (fast) launch_face_fast(c, sp.S);
    else if (c->prm.ischeme == 2) dispatch_face_fused(c, sp.S);
    else dispatch_face(c, sp.S);
  }
  mark(3);
  sp.Sf = fast ? c->Sfx : nullptr;
  if (!(tiles && c->gather_all_skip)) launch_gather(c, sp, write_rhs, true, first, irk, dt);
  if (tiles) {
    mark(4);
    fv_launch_tile_stage(c, sp.S, sp.SN, sp.Sout, sp.Sf, irk, dt, write_rhs, fast);
  }
  mark(-1);
  c->stage_fskip = c->stage_gskip = c->stage_cmask = nullptr;
  if (fast) std::swap(c->Sf, c->Sfx);
}

/* One time step (RK(0) + the method's stages), out of place: S^n stays in its
 * buffer and is the SN of every stage (no RK(0) copy is made), each stage writes
 * a buffer that is neither its input nor S^n, and the three state pointers are
 * then rotated so that c->S holds S^{n+1} and c->SN holds S^n, as after the
 * reference's RK(0) .. RK(3). */
static void time_step(fvg_ctx *c, int method, double dt, bool write_rhs_last, hipEvent_t *ev,
                      int *nev, int *kid) {
  int n;
  const int *st = stage_list(method, &n);
  const bool fast = fast_mixed(c);
  if (fast) ensure_sf(c);
  double *X0 = c->S, *buf[2] = {c->Sx, c->SN};
  const double *cur = X0;
  for (int k = 0; k < n; ++k) {
    double *out = buf[k & 1];
    stage(c, StagePtrs{cur, X0, out, nullptr}, st[k], dt, false, write_rhs_last && k == n - 1, ev,
          nev, kid, true);
    cur = out;
  }
  c->SN = X0;
  c->S = buf[(n - 1) & 1];
  c->Sx = buf[n & 1];
  /* the last stage read S^n itself (one stage) or the previous stage's output, which
   * the rotation leaves in Sx */
  c->stage_in = n == 1 ? c->SN : c->Sx;
  c->sf_valid = fast; /* the last gather wrote the FP32 copy of the new S */
}

/* ---- six steps as one HIP graph */

constexpr int kGraphSteps = 6; /* the state pointers cycle with period 3 (one or three
                                  stages per step) or 2 (two), the FP32 copies with 2 */

/* Everything the captured launches bake in: method, dt, the device pointers and sizes
 * they are given, the parameters and the tiling decisions. */
static void step_graph_key(const fvg_ctx *c, int method, double dt, std::vector<unsigned char> &k) {
  struct Fixed {
    int method, N, ns, F, nfp, Q, nlevel, ntile, nfa[3], nreg[3], usable, fall, gall, neq;
    size_t rsize;
    double dt;
    const void *ptr[23];
  } f{};
  f.method = method;
  f.N = c->N; f.ns = c->ns; f.F = c->F; f.nfp = c->nfp; f.Q = c->Q; f.nlevel = c->nlevel;
  f.ntile = c->ntile;
  for (int a = 0; a < 3; ++a) { f.nfa[a] = c->nfa[a]; f.nreg[a] = c->nreg[a]; }
  f.usable = fv_tiles_usable(c) ? 1 : 0;
  f.fall = c->face_all_skip ? 1 : 0;
  f.gall = c->gather_all_skip ? 1 : 0;
  f.neq = c->neq;
  f.rsize = c->rsize;
  f.dt = dt;
  const void *p[23] = {c->S, c->SN, c->Sx, c->Sf, c->Sfx, c->G, c->SGm, c->RHS, c->SF, c->err,
                       c->cflag, c->nbr, c->quad, c->fL, c->fR, c->fLL, c->fRR, c->fmeta, c->gslot,
                       c->tab, c->tiles, c->face_skip, c->tile_cell};
  for (int i = 0; i < 23; ++i) f.ptr[i] = p[i];
  k.resize(sizeof f + sizeof c->dp + sizeof c->prm + sizeof(void *));
  unsigned char *o = k.data();
  memcpy(o, &f, sizeof f);
  memcpy(o + sizeof f, &c->dp, sizeof c->dp);
  memcpy(o + sizeof f + sizeof c->dp, &c->prm, sizeof c->prm);
  const void *gs = c->gather_skip;
  memcpy(o + sizeof f + sizeof c->dp + sizeof c->prm, &gs, sizeof gs);
}

/* The graph of kGraphSteps steps for (method, dt), captured now if the cached one does
 * not match; false (eager launches) if capture fails. */
static bool step_graph_ready(fvg_ctx *c, int method, double dt) {
  if (fast_mixed(c)) ensure_sf(c); /* an eager conversion, not a node of the graph */
  std::vector<unsigned char> key;
  step_graph_key(c, method, dt, key);
  if (c->step_graph && key == c->step_graph_key) return true;
  release_step_graph(c);
  double *S0 = c->S, *SN0 = c->SN, *Sx0 = c->Sx, *stage_in0 = c->stage_in;
  float *Sf0 = c->Sf, *Sfx0 = c->Sfx;
  const bool sfv0 = c->sf_valid;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  for (int s = 0; s < kGraphSteps; ++s) time_step(c, method, dt, false, nullptr, nullptr, nullptr);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  const bool cycled = c->S == S0 && c->SN == SN0 && c->Sx == Sx0 && c->Sf == Sf0 && c->Sfx == Sfx0;
  /* nothing ran: the host's view of the buffers is the one before the capture */
  c->S = S0; c->SN = SN0; c->Sx = Sx0; c->Sf = Sf0; c->Sfx = Sfx0;
  c->stage_in = stage_in0;
  c->sf_valid = sfv0;
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess && cycled) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess || !cycled || !x) {
    if (x) (void)hipGraphExecDestroy(x);
    (void)hipGetLastError();
    return false;
  }
  c->step_graph = x;
  c->step_graph_key = std::move(key);
  return true;
}

/* nsteps time steps: whole groups of kGraphSteps as graph launches, the rest (always the
 * last step, which writes RHS if asked) eagerly */
static void run_steps(fvg_ctx *c, int nsteps, int method, double dt, bool write_rhs_last) {
  int body = c->graphs ? (nsteps - 1) / kGraphSteps * kGraphSteps : 0;
  if (body > 0 && !step_graph_ready(c, method, dt)) body = 0;
  for (int r = 0; r < body / kGraphSteps; ++r) {
    if (hipGraphLaunch(c->step_graph, c->stream) != hipSuccess) {
      (void)hipGetLastError();
      release_step_graph(c);
      body = r * kGraphSteps;
      break;
    }
  }
  if (body > 0) c->sf_valid = fast_mixed(c); /* the graph's last gather wrote the FP32 copy */
  for (int s = body; s < nsteps; ++s)
    time_step(c, method, dt, write_rhs_last && s == nsteps - 1, nullptr, nullptr, nullptr);
}

void fv_launch_recon(fvg_ctx *c, const double *S) { dispatch_recon(c, S); }

void fv_launch_grad_raw(fvg_ctx *c, const double *S) { dispatch_grad(c, S, false, true); }

void fv_launch_grad_tangent(fvg_ctx *c, const double *Sd, double *SGmd) {
  DevMesh M = dev_mesh(c);
  DevParams Pz = c->dp;
  for (int v = 0; v < 5; ++v) Pz.inlet[v] = 0.0;
  hipLaunchKernelGGL((k_grad<1, false, double, false, true>), dim3(nblocks(c->N)), dim3(FV_BLOCK), 0,
                     c->stream, M, Pz, Sd, (double *)nullptr, SGmd, (const uint8_t *)nullptr,
                     (const uint8_t *)nullptr);
}

/* the adjoint's forward stages: always the FP64-read path, whose face states
 * the adjoint face kernel rebuilds (recon_face) */
void fv_stage(fvg_ctx *c, int irk, double dt, bool first, bool write_rhs) {
  stage(c, StagePtrs{c->S, c->SN, c->S, nullptr}, irk, dt, first, write_rhs, nullptr, nullptr,
        nullptr, false);
  c->sf_valid = false;
  c->stage_in = nullptr;
}

void fv_stage_ptrs(fvg_ctx *c, const double *S, double *SN, double *Sout, int irk, double dt) {
  stage(c, StagePtrs{S, SN, Sout, nullptr}, irk, dt, false, false, nullptr, nullptr, nullptr, false);
  if (Sout == c->S) c->sf_valid = false;
  c->stage_in = nullptr;
}

void fv_launch_gather_ptrs(fvg_ctx *c, const double *G, double *S, double *SN, int irk, double dt,
                           bool first) {
  DevMesh M = dev_mesh(c);
  dim3 g(gather_grid(c->N)), b(FV_BLOCK);
  if (first)
    hipLaunchKernelGGL((k_gather_rk<5, double, false, true, true>), g, b, 0, c->stream, M, G, S, SN,
                       (double *)nullptr, irk, dt, c->err, S, (float *)nullptr,
                       (const uint8_t *)nullptr, (const uint8_t *)nullptr);
  else
    hipLaunchKernelGGL((k_gather_rk<5, double, false, true, false>), g, b, 0, c->stream, M, G, S, SN,
                       (double *)nullptr, irk, dt, c->err, S, (float *)nullptr,
                       (const uint8_t *)nullptr, (const uint8_t *)nullptr);
}

extern "C" {

int fvg_set_state(fvg_ctx *c, const double *S) { return copy_field(c, FVG_FIELD_S, (double *)S, true); }
int fvg_get_state(fvg_ctx *c, double *S) { return copy_field(c, FVG_FIELD_S, S, false); }
int fvg_set_field(fvg_ctx *c, int which, const double *src) {
  return copy_field(c, which, (double *)src, true);
}
int fvg_get_field(fvg_ctx *c, int which, double *dst) { return copy_field(c, which, dst, false); }

static int state_device_copy(fvg_ctx *c, void *d, bool to_ctx) {
  if (!c || !d) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  if (to_ctx) c->sf_valid = false;
  if (to_ctx) c->stage_in = nullptr;
  if (to_ctx)
    hipLaunchKernelGGL(k_transpose_state, dim3(nblocks(c->N)), dim3(FV_BLOCK), 0, c->stream,
                       (const double *)d, c->S, c->N, c->neq, 1);
  else
    hipLaunchKernelGGL(k_transpose_state, dim3(nblocks(c->N)), dim3(FV_BLOCK), 0, c->stream,
                       (const double *)c->S, (double *)d, c->N, c->neq, 0);
  int rc = finish(c, "state copy");
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

int fvg_set_state_device(fvg_ctx *c, const void *dS) { return state_device_copy(c, (void *)dS, true); }
int fvg_get_state_device(fvg_ctx *c, void *dS) { return state_device_copy(c, dS, false); }

int fvg_device_ptr(fvg_ctx *c, int which, void **ptr) {
  if (!c || !ptr) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  int len;
  bool isR;
  if (!field_rows(c, which, &len, ptr, &isR)) return set_err(c, FVG_ERR_ARG, "unknown field");
  return FVG_OK;
}

int fvg_faceinterp(fvg_ctx *c) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  (void)hipSetDevice(c->device);
  dispatch_recon(c, c->S);
  return finish(c, "faceinterp");
}

int fvg_gradcalc(fvg_ctx *c) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  (void)hipSetDevice(c->device);
  dispatch_grad(c, c->S, true);
  if (c->prm.ns) dispatch_grad(c, c->S, false);
  return finish(c, "gradcalc");
}

/* gradcalc, then the flags of every cell into the device buffer d (>= N bytes), then
 * to the host */
static int adapt_flags_via(fvg_ctx *c, int kind, double refine, double coarsen, int base_level,
                           int max_level, int8_t *d, int8_t *flags) {
  dispatch_grad(c, c->S, true);
  DevMesh M = dev_mesh(c);
  hipLaunchKernelGGL(k_adapt_flags, dim3(nblocks(c->N)), dim3(FV_BLOCK), 0, c->stream, M,
                     (const double *)c->S, c->neq, (const double *)c->SG, kind, refine, coarsen,
                     base_level, max_level, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(flags, d, (size_t)c->N, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  return FVG_OK;
}

int fvg_adapt_flags(fvg_ctx *c, int kind, double refine, double coarsen, int base_level,
                    int max_level, int8_t *flags) {
  if (!c || !flags) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  if (kind != 0 && kind != 1) return set_err(c, FVG_ERR_ARG, "kind must be 0 or 1");
  (void)hipSetDevice(c->device);
  int8_t *d = nullptr;
  HIPCHK(c, hipMalloc((void **)&d, (size_t)c->ns));
  int rc = adapt_flags_via(c, kind, refine, coarsen, base_level, max_level, d, flags);
  (void)hipFree(d);
  return rc;
}

int fvg_adapt_flags_static(fvg_ctx *c, int kind, double refine, double coarsen, int base_level,
                           int max_level, int8_t *flags) {
  if (!c || !flags) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  if (kind != 0 && kind != 1) return set_err(c, FVG_ERR_ARG, "kind must be 0 or 1");
  if (c->N > FVG_STATIC_FLAGS_MAX) {
    char b[160];
    snprintf(b, sizeof b, "%d cells exceed the static flag buffer of %d cells (FVG_STATIC_FLAGS_MAX)",
             c->N, (int)FVG_STATIC_FLAGS_MAX);
    return set_err(c, FVG_ERR_ARG, b);
  }
  if (c->flag_scratch_rc != (int)hipSuccess || !c->flag_scratch)
    return set_err(c, FVG_ERR_HIP, "no device address for the static flag buffer");
  (void)hipSetDevice(c->device);
  return adapt_flags_via(c, kind, refine, coarsen, base_level, max_level, c->flag_scratch, flags);
}

int fvg_pressure(fvg_ctx *c, double *p) {
  if (!c || !p) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  double *d = nullptr;
  HIPCHK(c, hipMalloc((void **)&d, (size_t)c->N * sizeof(double)));
  DevMesh M = dev_mesh(c);
  dim3 g(nblocks(c->N)), b(FV_BLOCK);
  switch (c->prm.model) {
    case 1: hipLaunchKernelGGL(k_pressure<1>, g, b, 0, c->stream, M, c->dp, (const double *)c->S, d); break;
    case 2: hipLaunchKernelGGL(k_pressure<2>, g, b, 0, c->stream, M, c->dp, (const double *)c->S, d); break;
    default: hipLaunchKernelGGL(k_pressure<3>, g, b, 0, c->stream, M, c->dp, (const double *)c->S, d); break;
  }
  hipError_t e = hipMemcpyAsync(p, d, (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  return finish(c, "pressure");
}

int fvg_rhs(fvg_ctx *c) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  (void)hipSetDevice(c->device);
  reset_err(c);
  dispatch_face(c, c->S);
  launch_gather(c, true, false, false, 0, 0.0);
  int rc = finish(c, "rhs");
  if (rc) return rc;
  int32_t bad;
  rc = read_err(c, &bad);
  if (rc) return rc;
  if (bad >= 0) {
    char b[96];
    snprintf(b, sizeof b, "non-finite residual or failed Riemann/EOS at cell %d", bad);
    return set_err(c, FVG_ERR_NONFINITE, b);
  }
  return FVG_OK;
}

int fvg_rk(fvg_ctx *c, int irk, double dt) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  if (irk < 0 || irk > 4) return set_err(c, FVG_ERR_ARG, "irk must be 0..4");
  (void)hipSetDevice(c->device);
  DevMesh M = dev_mesh(c);
  dim3 g(nblocks(c->N)), b(FV_BLOCK);
  c->sf_valid = false;
  c->stage_in = nullptr;
  if (c->neq == 5)
    hipLaunchKernelGGL(k_rk<5>, g, b, 0, c->stream, M, c->S, c->SN, c->RHS, irk, dt);
  else
    hipLaunchKernelGGL(k_rk<4>, g, b, 0, c->stream, M, c->S, c->SN, c->RHS, irk, dt);
  return finish(c, "rk");
}

int fvg_steps(fvg_ctx *c, int nsteps, int method, double dt) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  if (method < 1 || method > 3 || nsteps < 0) return set_err(c, FVG_ERR_ARG, "bad method/nsteps");
  (void)hipSetDevice(c->device);
  reset_err(c);
  /* RHS is written on the last stage only: afterwards it holds that stage's residual */
  run_steps(c, nsteps, method, dt, true);
  int rc = finish(c, "steps");
  if (rc) return rc;
  int32_t bad;
  rc = read_err(c, &bad);
  if (rc) return rc;
  if (bad >= 0) {
    char b[96];
    snprintf(b, sizeof b, "non-finite residual or failed Riemann/EOS at cell %d", bad);
    return set_err(c, FVG_ERR_NONFINITE, b);
  }
  return FVG_OK;
}

int fvg_step(fvg_ctx *c, int method, double dt) { return fvg_steps(c, 1, method, dt); }

int fvg_time_steps(fvg_ctx *c, int nsteps, int method, double dt, double *ms_total,
                   double *ms_kernel) {
  if (!c || !c->has_mesh) return c ? set_err(c, FVG_ERR_STATE, "no mesh") : FVG_ERR_ARG;
  if (method < 1 || method > 3 || nsteps < 1) return set_err(c, FVG_ERR_ARG, "bad method/nsteps");
  (void)hipSetDevice(c->device);
  if (!c->ev_ok) {
    for (auto &e : c->ev) HIPCHK(c, hipEventCreate(&e));
    c->ev_ok = true;
  }
  reset_err(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  hipEvent_t &t0 = c->ev[2 * 64], &t1 = c->ev[2 * 64 + 1];
  HIPCHK(c, hipEventRecord(t0, c->stream));
  run_steps(c, nsteps, method, dt, false);
  HIPCHK(c, hipEventRecord(t1, c->stream));
  HIPCHK(c, hipEventSynchronize(t1));
  float tot = 0.f;
  HIPCHK(c, hipEventElapsedTime(&tot, t0, t1));
  if (ms_total) *ms_total = tot;
  if (ms_kernel) {
    /* Mean duration of each kernel of a stage, from batches of kReps back-to-back
     * launches between two events (an event pair around every launch would add the
     * launch latency to each).  Every launch repeats the same work: the gradient and
     * face kernels rewrite SGm and G from the current S, the gather writes the spare
     * buffer (its FP32 copy, in the mixed form, into the unused SF array).  The
     * gather's mean weighs the first stage (SN = S, read once) 1 : 2 against the
     * others, as in a step of the three-stage method. */
    constexpr int kReps = 10;
    const bool fast = fast_mixed(c);
    if (fast) ensure_sf(c);
    c->stage_in = nullptr; /* the batches below overwrite the spare buffer */
    bool ev_fail = false;
    auto batch = [&](auto &&launch) -> double {
      float ms = 0.f;
      ev_fail |= hipEventRecord(c->ev[0], c->stream) != hipSuccess;
      for (int r = 0; r < kReps; ++r) launch();
      ev_fail |= hipEventRecord(c->ev[1], c->stream) != hipSuccess;
      ev_fail |= hipEventSynchronize(c->ev[1]) != hipSuccess;
      ev_fail |= hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) != hipSuccess;
      return (double)ms / kReps;
    };
    for (int k = 0; k < FV_NKT; ++k) ms_kernel[k] = 0.0;
    float *sf_dummy = fast ? (float *)c->SF : nullptr;
    const bool tiles = fv_tiles_usable(c);
    c->stage_fskip = tiles ? c->face_skip : nullptr;
    c->stage_gskip = tiles ? c->gather_skip : nullptr;
    c->stage_cmask = tiles ? c->tile_cell : nullptr;
    const float *sfin = fast ? c->Sf : nullptr;
    if (c->prm.ns && !(tiles && c->gather_all_skip))
      ms_kernel[1] = batch([&] { dispatch_grad(c, c->S, false, false, sfin); });
    if (c->prm.ns && tiles) ms_kernel[5] = batch([&] { fv_launch_tile_grad(c, c->S, sfin); });
    if (!(tiles && c->face_all_skip))
      ms_kernel[2] = batch([&] {
        if (fast) launch_face_fast(c, c->S);
        else if (c->prm.ischeme == 2) dispatch_face_fused(c, c->S);
        else dispatch_face(c, c->S);
      });
    if (!(tiles && c->gather_all_skip)) {
      const double g1 = batch([&] {
        launch_gather(c, StagePtrs{c->S, c->S, c->Sx, sf_dummy}, false, true, false, 1, dt);
      });
      const double g2 = batch([&] {
        launch_gather(c, StagePtrs{c->S, c->SN, c->Sx, sf_dummy}, false, true, false, 2, dt);
      });
      ms_kernel[3] = (g1 + 2.0 * g2) / 3.0;
    }
    if (tiles)
      ms_kernel[4] = batch([&] {
        fv_launch_tile_stage(c, c->S, c->SN, c->Sx, sf_dummy, 2, dt, false, fast);
      });
    c->stage_fskip = c->stage_gskip = c->stage_cmask = nullptr;
    if (ev_fail) return set_err(c, FVG_ERR_HIP, "kernel timing events failed");
  }
  int32_t bad;
  int rc = read_err(c, &bad);
  if (rc) return rc;
  if (bad >= 0) return set_err(c, FVG_ERR_NONFINITE, "non-finite state during timed steps");
  return FVG_OK;
}

int fvg_eos_host(const fvg_params *p, const double *U, double *out) {
  if (!p || !U || !out || fvg_neq(p->model) < 0) return FVG_ERR_ARG;
  DevParams P;
  memset(&P, 0, sizeof P);
  to_dev_params(p, &P);
  double r, u[3], e, pr = 0.0, c = 0.0, y;
  bool ok;
  switch (p->model) {
    case 1: ok = eos64<1>(P, U, false, r, u, e, pr, c, y); break;
    case 2: ok = eos64<2>(P, U, false, r, u, e, pr, c, y); break;
    default: ok = eos64<3>(P, U, false, r, u, e, pr, c, y); break;
  }
  out[0] = pr;
  out[1] = c;
  return ok ? FVG_OK : FVG_ERR_NONFINITE;
}

int fvg_cfl_dt(fvg_ctx *c, double cfl, double *dt) {
  if (!c || !dt) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  reset_err(c);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->redu, 0ull);
  DevMesh M = dev_mesh(c);
  dim3 g(nblocks(c->N)), b(FV_BLOCK);
  switch (c->prm.model) {
    case 1: hipLaunchKernelGGL(k_cfl<1>, g, b, 0, c->stream, M, c->dp, c->S, c->redu, c->err); break;
    case 2: hipLaunchKernelGGL(k_cfl<2>, g, b, 0, c->stream, M, c->dp, c->S, c->redu, c->err); break;
    default: hipLaunchKernelGGL(k_cfl<3>, g, b, 0, c->stream, M, c->dp, c->S, c->redu, c->err); break;
  }
  int rc = finish(c, "cfl");
  if (rc) return rc;
  unsigned long long bits;
  HIPCHK(c, hipMemcpyAsync(&bits, c->redu, 8, hipMemcpyDeviceToHost, c->stream));
  int32_t bad;
  rc = read_err(c, &bad);
  if (rc) return rc;
  if (bad >= 0) return set_err(c, FVG_ERR_NONFINITE, "non-physical state in CFL scan");
  double m;
  memcpy(&m, &bits, 8);
  if (!(m > 0.0)) return set_err(c, FVG_ERR_NONFINITE, "zero wave speed");
  *dt = cfl / m;
  return FVG_OK;
}

int fvg_totals(fvg_ctx *c, double *out) {
  if (!c || !out) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  const int chunk = 16 * FV_BLOCK;
  const int nb = (c->N + chunk - 1) / chunk;
  const size_t need = (size_t)(nb + 1) * (c->neq + 1);
  if (c->red_cap < need) {
    if (c->red) (void)hipFree(c->red);
    c->red = nullptr;
    HIPCHK(c, hipMalloc(&c->red, need * 8));
    c->red_cap = need;
  }
  DevMesh M = dev_mesh(c);
  double *fin = c->red + (size_t)nb * (c->neq + 1);
  if (c->neq == 5) {
    hipLaunchKernelGGL(k_totals_part<5>, dim3(nb), dim3(FV_BLOCK), 0, c->stream, M, c->S, c->red, chunk);
    hipLaunchKernelGGL(k_totals_final<5>, dim3(1), dim3(64), 0, c->stream, c->red, nb, fin);
  } else {
    hipLaunchKernelGGL(k_totals_part<4>, dim3(nb), dim3(FV_BLOCK), 0, c->stream, M, c->S, c->red, chunk);
    hipLaunchKernelGGL(k_totals_final<4>, dim3(1), dim3(64), 0, c->stream, c->red, nb, fin);
  }
  int rc = finish(c, "totals");
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, fin, (size_t)(c->neq + 1) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

int fvg_check_state(fvg_ctx *c, int32_t *bad) {
  if (!c || !bad) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  reset_err(c);
  DevMesh M = dev_mesh(c);
  if (c->neq == 5)
    hipLaunchKernelGGL(k_check<5>, dim3(nblocks(c->N)), dim3(FV_BLOCK), 0, c->stream, M, c->S, c->RHS, c->err);
  else
    hipLaunchKernelGGL(k_check<4>, dim3(nblocks(c->N)), dim3(FV_BLOCK), 0, c->stream, M, c->S, c->RHS, c->err);
  int rc = finish(c, "check");
  if (rc) return rc;
  return read_err(c, bad);
}

int fvg_sync(fvg_ctx *c) {
  if (!c) return FVG_ERR_ARG;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVG_OK;
}

}  // extern "C"

/* ------------------------------------------------------------------ bandwidth probe */

/* A plain streaming copy: each thread moves four 16-byte words, all four loads issued
 * before the stores, consecutive lanes on consecutive words (1 KiB per wave and load). */
constexpr int kCopyBlock = 256, kCopyPerThread = 4;

__global__ __launch_bounds__(kCopyBlock) void k_stream_copy(const double2 *__restrict__ s,
                                                            double2 *__restrict__ d, size_t n) {
  const size_t base = (size_t)blockIdx.x * (kCopyBlock * kCopyPerThread) + threadIdx.x;
  double2 v[kCopyPerThread];
#pragma unroll
  for (int k = 0; k < kCopyPerThread; ++k) {
    const size_t i = base + (size_t)k * kCopyBlock;
    if (i < n) v[k] = s[i];
  }
#pragma unroll
  for (int k = 0; k < kCopyPerThread; ++k) {
    const size_t i = base + (size_t)k * kCopyBlock;
    if (i < n) d[i] = v[k];
  }
}

extern "C" int fvg_stream_copy(int device, void *dst, const void *src, size_t bytes, int reps,
                               double *ms) {
  if (!dst || !src || !ms || reps < 1 || bytes == 0 || bytes % 16 ||
      ((uintptr_t)dst | (uintptr_t)src) % 16)
    return FVG_ERR_ARG;
  const size_t n = bytes / 16;
  const size_t per_block = (size_t)kCopyBlock * kCopyPerThread;
  const size_t blocks = (n + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return FVG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return FVG_ERR_HIP;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) return FVG_ERR_HIP;
  int rc = FVG_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = FVG_ERR_HIP;
  if (!rc) {
    const dim3 g((unsigned)blocks), b(kCopyBlock);
    hipLaunchKernelGGL(k_stream_copy, g, b, 0, st, (const double2 *)src, (double2 *)dst, n);
    bool ok = hipEventRecord(e0, st) == hipSuccess;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, g, b, 0, st, (const double2 *)src, (double2 *)dst, n);
    ok = ok && hipGetLastError() == hipSuccess;
    ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0.f;
    ok = ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
    if (ok) *ms = (double)t / reps;
    else rc = FVG_ERR_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return rc;
}
