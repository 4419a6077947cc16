/*
 * LDS-tiled RK stage for the uniform parts of a mesh.
 *
 * A tile is an 8 x 8 x 8 block of cells of one level that are consecutive in the leaf
 * list (Morton order: the local index interleaves the bits of x, y, z with x lowest, as
 * the depth-first order over kids by corner node gives, split.c:124-128); an interior
 * tile is one whose six face-neighbour blocks are tiles too.  One workgroup of 512
 * threads (one per cell) does a whole stage for an interior tile (k_tile_stage):
 *   1. loads the tile from HBM into LDS once and, per axis, the two cell layers on either
 *      side of it along that axis (the MUSCL stencils of the axis's faces);
 *   2. per axis, evaluates each face of the tile once from LDS: the regular-face MUSCL
 *      of reconctruct_V3.c, the EOS, HLLC and (Navier-Stokes) the viscous flux with the
 *      viscous rows k_tile_grad / k_grad wrote (three per cell for an axis's faces, four
 *      with heat conduction); the face's flux goes to an LDS buffer;
 *   3. sums each cell's six face fluxes in the gather's order (-x, +x, -y, +y, -z, +z;
 *      residual_faces_AMR.c:173-175) and applies the RK stage (RK.c:24-47).
 * k_tile_grad forms the Green-Gauss gradients of the tiles' cells from the tile and the
 * cell layer beyond each face, staged in LDS, with k_grad's arithmetic, and stores their
 * viscous rows (vrows) as k_grad does (both read the FP32 copy of the state in the mixed
 * ideal-gas form, whose viscous flux is FP32).
 * Face fluxes never go to HBM and every stencil row is read once per tile, where the
 * face-list path writes the face flux, reads it back from both cells and fetches each
 * stencil row for every face that uses it.  A face is evaluated exactly as the face-list
 * kernel evaluates it (the same orientation: the lower-indexed cell is the owner), so a
 * tile cell's new state is bitwise the one the face-list path gives.
 *
 * The rest of the mesh (tiles at walls or at level changes, hanging faces, anything
 * that is not a tile) stays on the face-list kernels, which skip the blocks whose faces
 * and cells are all the interior tiles' (and, in a block of cells covered in part, those
 * cells).
 */
#include <type_traits>

#include "fv_common.h"

namespace {

constexpr int TT = 512;          /* cells (threads) of a tile */
constexpr int TREG = 512 + 256;  /* the tile + the two slabs (2 x 8 x 8 cells) of one axis */
constexpr int TFB = 576;         /* faces of one axis: 9 x 8 x 8 */

/* local index of cell (x, y, z) of a tile: bits x0 y0 z0 x1 y1 z1 x2 y2 z2 */
__host__ __device__ __forceinline__ int tmorton(int x, int y, int z) {
  return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) |
         ((z & 2) << 4) | ((x & 4) << 4) | ((y & 4) << 5) | ((z & 4) << 6);
}

__host__ __device__ __forceinline__ void demorton(int t, int &x, int &y, int &z) {
  x = (t & 1) | ((t >> 2) & 2) | ((t >> 4) & 4);
  y = ((t >> 1) & 1) | ((t >> 3) & 2) | ((t >> 5) & 4);
  z = ((t >> 2) & 1) | ((t >> 4) & 2) | ((t >> 6) & 4);
}

/* Tiles are listed in leaf (Morton) order and taken in the XCD-contiguous order of
 * logical_block(): each XCD works through one run of neighbouring tiles, whose slabs
 * and gradient rows then come from its own L2. */

/* a tile's first cell and those of its neighbour tiles (a row of the tile table:
 * base, then across faces 0 -y, 1 +x, 2 +y, 3 -x, 4 -z, 5 +z, then the level) */
struct TileLinks {
  int base, my, px, py, mx, mz, pz;
};

/* cell coordinates of position p along axis a, the other two coordinates u, w (in
 * axis order) */
template <int a>
__device__ __forceinline__ void along(int p, int u, int w, int &x, int &y, int &z) {
  if constexpr (a == 0) { x = p; y = u; z = w; }
  else if constexpr (a == 1) { x = u; y = p; z = w; }
  else { x = u; y = w; z = p; }
}

/* Region during the pass of axis a: the tile (slot = local index) and the two cell
 * layers on either side along a (slot 512 + side x 128 + depth x 64 + u x 8 + w; side 0
 * below the tile, depth 0 next to it).  p in [-2, 10). */
template <int a>
__device__ __forceinline__ int reg_slot(int p, int u, int w) {
  if (p < 0) return 512 + (-1 - p) * 64 + u * 8 + w;
  if (p > 7) return 640 + (p - 8) * 64 + u * 8 + w;
  int x, y, z;
  along<a>(p, u, w, x, y, z);
  return tmorton(x, y, z);
}

/* the cell at position p of the row u, w along axis a: the tile's, or the neighbour
 * tile's below / above it.  The first cell is picked by value (indexing an array of the
 * seven would put the array in scratch memory). */
template <int a>
__device__ __forceinline__ int reg_cell(const TileLinks &tb, int p, int u, int w) {
  int b = tb.base;
  if (p < 0) { b = a == 0 ? tb.mx : (a == 1 ? tb.my : tb.mz); p += 8; }
  else if (p > 7) { b = a == 0 ? tb.px : (a == 1 ? tb.py : tb.pz); p -= 8; }
  int x, y, z;
  along<a>(p, u, w, x, y, z);
  return b + tmorton(x, y, z);
}

/* The cells and region slots of one face of axis a, between the cells at positions
 * p - 1 and p of a row: cm, cp the two cells, c0, c3 those at p - 2 and p + 1 (read only
 * by the mixed form's FP64 rebuild), r0..r3 the slots of positions p - 2 .. p + 1. */
struct FaceIdx {
  int cm, cp, c0, c3, r0, r1, r2, r3;
};

/* those of the face at position p (in [0, 8]) of the row u, w */
template <int a>
__device__ __forceinline__ FaceIdx face_idx(const TileLinks &tb, int p, int u, int w) {
  return FaceIdx{reg_cell<a>(tb, p - 1, u, w), reg_cell<a>(tb, p, u, w),
                 reg_cell<a>(tb, p - 2, u, w), reg_cell<a>(tb, p + 1, u, w),
                 reg_slot<a>(p - 2, u, w), reg_slot<a>(p - 1, u, w), reg_slot<a>(p, u, w),
                 reg_slot<a>(p + 1, u, w)};
}

/* the same for the -a face of the tile's cell t at position c along a (uw = u x 8 + w),
 * stepping along a in the local Morton index: the bits of axis a sit at a, a + 3, a + 6
 * (mask ma); with the other bits o = t & ~ma, the cell at position q of the row is
 * o | spread(q), and spread(q -+ 1) = (spread(q) -+ 1 in those bits) & ma, a step from 0
 * down wrapping to 7 (the neighbour tile's last layer) */
template <int a>
__device__ __forceinline__ FaceIdx own_face_idx(const TileLinks &tb, int t, int c, int uw) {
  constexpr int ma = 0x49 << a, one = 1 << a;
  const int o = t & ~ma, s0 = t & ma;
  const int sm1 = (s0 - one) & ma, sm2 = (sm1 - one) & ma, sp1 = ((s0 | ~ma) + one) & ma;
  const int bm = a == 0 ? tb.mx : (a == 1 ? tb.my : tb.mz);
  const int bp = a == 0 ? tb.px : (a == 1 ? tb.py : tb.pz);
  FaceIdx f;
  f.cm = (c > 0 ? tb.base : bm) + (o | sm1);
  f.cp = tb.base + t;
  f.c0 = (c > 1 ? tb.base : bm) + (o | sm2);
  f.c3 = (c < 7 ? tb.base : bp) + (o | sp1);
  f.r0 = c > 1 ? (o | sm2) : 512 + (1 - c) * 64 + uw;
  f.r1 = c > 0 ? (o | sm1) : 512 + uw;
  f.r2 = t;
  f.r3 = c < 7 ? (o | sp1) : 640 + uw;
  return f;
}

/* One face of axis a (cells and slots fi).  The owner is the lower-indexed of the two
 * cells, as in the face list: the flux is the face-list kernel's for that face
 * (recon_regular, make_state, hllc_face, -Area), returned as the -1 side cell sees it
 * through its +a face. */
template <int MODEL, bool NS, int a, typename T, typename R, int OPT, bool FAST, typename ST>
__device__ __forceinline__ void tile_face(const DevMesh &M, const DevParams &P, const TileLinks &tb,
                                          int L, const ST *reg, const double *__restrict__ S,
                                          const R *__restrict__ SGm, const FaceIdx &fi, R *out,
                                          int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int cm = fi.cm, cp = fi.cp;
  const bool own_minus = cm < cp;
  /* stencil in the owner's orientation: far side of the owner, owner, other, far side */
  const int q0 = own_minus ? fi.r0 : fi.r3, q1 = own_minus ? fi.r1 : fi.r2,
            q2 = own_minus ? fi.r2 : fi.r1, q3 = own_minus ? fi.r3 : fi.r0;
  const int iL = own_minus ? cm : cp, iR = own_minus ? cp : cm;
  const T s = own_minus ? (T)1 : (T)-1;
  const int k = (L * 3 + a) * 3;
  const double pcf = M.dist[2 * k], ic = M.idist[k];
  FState<T> Ls, Rs;
  bool ok = true;
  bool done = false;
  if constexpr (FAST) {
    /* the FP32 face states of the mixed ideal-gas kernel (face_fast_body, regular) */
    const float fp = (float)pcf, fic = (float)ic, fpic = fp * fic;
    float U[5], V[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const float A0 = reg[v * TREG + q0], A1 = reg[v * TREG + q1];
      const float B0 = reg[v * TREG + q2], B1 = reg[v * TREG + q3];
      U[v] = muscl_regular_f(P.limiter, P.va_eps, A1, B0, A0, fp, fic, fpic);
      V[v] = muscl_regular_f(P.limiter, P.va_eps, B0, A1, B1, fp, fic, fpic);
    }
    done = make_state_mixed_ideal(P, U, Ls);
    done = make_state_mixed_ideal(P, V, Rs) && done;
  }
  if (!done) {
    /* FP64 stencil values: from LDS, or (the FP32 inversion rejected a state) the FP64
     * rebuild of face_rebuild_block from HBM */
    double UL[NEQ], UR[NEQ];
    const double *g0, *g1, *g2, *g3;
    if constexpr (FAST) {
      const int c0 = fi.c0, c3 = fi.c3;
      g0 = S + (size_t)(own_minus ? c0 : c3) * NEQ;
      g1 = S + (size_t)iL * NEQ;
      g2 = S + (size_t)iR * NEQ;
      g3 = S + (size_t)(own_minus ? c3 : c0) * NEQ;
    }
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      double A0, A1, B0, B1;
      if constexpr (FAST) {
        A0 = g0[v]; A1 = g1[v]; B0 = g2[v]; B1 = g3[v];
      } else {
        A0 = reg[v * TREG + q0]; A1 = reg[v * TREG + q1];
        B0 = reg[v * TREG + q2]; B1 = reg[v * TREG + q3];
      }
      const T gL = limited_regular<T>(P.limiter, P.va_eps, (T)(B0 - A1), (T)(A0 - A1), (T)ic);
      const T gR = limited_regular<T>(P.limiter, P.va_eps, (T)(A1 - B0), (T)(B1 - B0), (T)ic);
      UL[v] = A1 + (double)((T)pcf * gL);
      UR[v] = B0 + (double)((T)pcf * gR);
    }
    ok = make_state<MODEL, T>(P, UL, true, Ls);
    ok = make_state<MODEL, T>(P, UR, true, Rs) && ok;
  }
  if constexpr (NS) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ls.tv[k] = (T)SGm[(size_t)vrow(a, k) * M.ns + iL];
      Rs.tv[k] = (T)SGm[(size_t)vrow(a, k) * M.ns + iR];
    }
    if constexpr ((OPT & OPT_HEAT) != 0) {
      Ls.gt[a] = (T)SGm[(size_t)(VROW_HEAT + a) * M.ns + iL];
      Rs.gt[a] = (T)SGm[(size_t)(VROW_HEAT + a) * M.ns + iR];
    }
  }
  T o[NEQ];
  ok = hllc_face<MODEL, NS, a, T, OPT>(P, Ls, Rs, s, o) && ok;
  const T A = (T)M.area[3 * L + a];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const R g = (R)(-(A) * o[v]);
    out[v] = own_minus ? g : -g;
  }
  if (!ok) atomicMin(err, iL);
}

/* the two slabs of axis a (256 cells, one per thread t < 256): fetched into registers,
 * then stored into the region's slab slots */
template <int a, bool FAST, int NEQ, typename ST>
__device__ __forceinline__ void fetch_slabs(const TileLinks &tb, const double *__restrict__ S,
                                            const float *__restrict__ Sf, ST *val, int t) {
  if (t < 256) {
    const int side = t >> 7, d = (t >> 6) & 1, u = (t >> 3) & 7, w = t & 7;
    const int cell = reg_cell<a>(tb, side ? 8 + d : -1 - d, u, w);
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      if constexpr (FAST) val[v] = Sf[(size_t)cell * NEQ + v];
      else val[v] = S[(size_t)cell * NEQ + v];
    }
  }
}

template <int NEQ, typename ST>
__device__ __forceinline__ void store_slabs(const ST *val, ST *reg, int t) {
  if (t < 256) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) reg[v * TREG + 512 + t] = val[v];
  }
}

template <int a, bool FAST, int NEQ, typename ST>
__device__ __forceinline__ void load_slabs(const TileLinks &tb, const double *__restrict__ S,
                                           const float *__restrict__ Sf, ST *reg, int t) {
  ST val[NEQ];
  fetch_slabs<a, FAST, NEQ, ST>(tb, S, Sf, val, t);
  store_slabs<NEQ, ST>(val, reg, t);
}

/* One axis: the slabs of a are in the region.  The -a face of every cell (fb slot = its
 * local index) and the +a faces on the tile's +a side (slots 512 + u x 8 + w, by the
 * first wave) into fb; then each cell's -a and +a terms into its sum acc (the gather's
 * terms: 0 + the signed flux, added to the sum). */
template <int MODEL, bool NS, int a, int NA, typename T, typename R, int OPT, bool FAST, typename ST>
__device__ __forceinline__ void tile_pass(const DevMesh &M, const DevParams &P, const TileLinks &tb,
                                          int L, ST *reg, R *fb, double *acc,
                                          const double *__restrict__ S, const float *__restrict__ Sf,
                                          const R *__restrict__ SGm,
                                          int t, int cx, int cy, int cz, int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const int c = a == 0 ? cx : (a == 1 ? cy : cz);
  const int u = a == 0 ? cy : cx, w = a == 2 ? cy : cz;
  /* the next axis's slabs (NA >= 0) are fetched before this axis's faces and stored once
   * they are done, so that their loads are in flight during the face arithmetic */
  ST nx[NEQ];
  if constexpr (NA >= 0) fetch_slabs<NA, FAST, NEQ, ST>(tb, S, Sf, nx, t);
  R out[NEQ];
  tile_face<MODEL, NS, a, T, R, OPT, FAST, ST>(M, P, tb, L, reg, S, SGm,
                                               own_face_idx<a>(tb, t, c, u * 8 + w), out, err);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) fb[v * TFB + t] = out[v];
  if (t < 64) {
    tile_face<MODEL, NS, a, T, R, OPT, FAST, ST>(M, P, tb, L, reg, S, SGm,
                                                 face_idx<a>(tb, 8, t >> 3, t & 7), out, err);
#pragma unroll
    for (int v = 0; v < NEQ; ++v) fb[v * TFB + 512 + t] = out[v];
  }
  __syncthreads();
  if constexpr (NA >= 0) store_slabs<NEQ, ST>(nx, reg, t);
  const int sp = c < 7 ? tmorton(cx + (a == 0), cy + (a == 1), cz + (a == 2)) : 512 + u * 8 + w;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    double m = 0.0, q = 0.0;
    m += -(double)fb[v * TFB + t];
    q += (double)fb[v * TFB + sp];
    double r = acc[v * TT + t];
    r += m;
    r += q;
    acc[v * TT + t] = r;
  }
}

/* FP64: four waves per SIMD (128 VGPRs), two workgroups per CU (74 KB of LDS each), so
 * that one's loads and barriers overlap the other's face arithmetic; the running sums
 * stay in LDS, not in registers, which is what lets the viscous FP64 faces fit that
 * budget.  FP32 reads (mixed): six waves per SIMD (80 VGPRs), three workgroups per CU
 * (47 KB each), measured 1.2x faster than four. */
template <int MODEL, bool NS, typename T, typename R, int OPT, bool FAST>
__global__ __launch_bounds__(TT, FAST ? 6 : 4) void k_tile_stage(DevMesh M, DevParams P,
                                                   const int32_t *__restrict__ tiles,
                                                   const double *__restrict__ S,
                                                   const float *__restrict__ Sf,
                                                   const double *__restrict__ SN,
                                                   const R *__restrict__ SGm, int irk, double dt,
                                                   double *__restrict__ Sout,
                                                   float *__restrict__ SfOut,
                                                   double *__restrict__ RHS, int write_rhs,
                                                   int *__restrict__ err) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  using ST = typename std::conditional<FAST, float, double>::type;
  /* one LDS buffer: region [NEQ][TREG] (ST), face fluxes [NEQ][TFB] (R), sums [NEQ][TT]
   * (double) */
  constexpr int B_REG = (int)sizeof(ST) * NEQ * TREG, B_FB = (int)sizeof(R) * NEQ * TFB;
  static_assert(B_REG % 16 == 0 && B_FB % 16 == 0, "LDS carve alignment");
  __shared__ __align__(16) unsigned char smem[B_REG + B_FB + 8 * NEQ * TT];
  ST *reg = (ST *)smem;
  R *fb = (R *)(smem + B_REG);
  double *acc = (double *)(smem + B_REG + B_FB);
  const int t = threadIdx.x;
  const int32_t *row = tiles + (size_t)logical_block() * 8;
  const TileLinks tb{row[0], row[1], row[2], row[3], row[4], row[5], row[6]};
  const int L = row[7];
  /* 1. the tile (its cells are consecutive: coalesced rows) and the x slabs into LDS */
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    if constexpr (FAST) reg[v * TREG + t] = Sf[(size_t)(tb.base + t) * NEQ + v];
    else reg[v * TREG + t] = S[(size_t)(tb.base + t) * NEQ + v];
    acc[v * TT + t] = 0.0;
  }
  load_slabs<0, FAST, NEQ, ST>(tb, S, Sf, reg, t);
  __syncthreads();
  int cx, cy, cz;
  demorton(t, cx, cy, cz);
  /* 2. faces axis by axis, each cell's sum in the gather's order (-x, +x, -y, +y, -z,
   * +z); the next axis's slabs replace this one's once its faces are done */
  tile_pass<MODEL, NS, 0, 1, T, R, OPT, FAST, ST>(M, P, tb, L, reg, fb, acc, S, Sf, SGm, t, cx, cy, cz, err);
  __syncthreads();
  tile_pass<MODEL, NS, 1, 2, T, R, OPT, FAST, ST>(M, P, tb, L, reg, fb, acc, S, Sf, SGm, t, cx, cy, cz, err);
  __syncthreads();
  /* FP64 (the mixed kernel has no registers to spare for it): the update's SN row is
   * fetched before the last pass, in the registers the slab prefetch of the other passes
   * uses, so that its loads are in flight during it */
  const size_t crow = (size_t)(tb.base + t) * NEQ;
  double sn[NEQ];
  if constexpr (!FAST) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) sn[v] = SN[crow + v];
  }
  tile_pass<MODEL, NS, 2, -1, T, R, OPT, FAST, ST>(M, P, tb, L, reg, fb, acc, S, Sf, SGm, t, cx, cy, cz, err);
  double rhs[NEQ];
  bool bad = false;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    rhs[v] = acc[v * TT + t];
    bad |= (rhs[v] != rhs[v]);
  }
  if (bad) atomicMin(err, tb.base + t);
  /* 3. the stage update of the tile's 512 consecutive cells: each thread its own cell's
   * rows of S, SN and Sout (a wave's five loads of a row cover one contiguous span) */
  double own[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    if constexpr (FAST) {
      own[v] = S[crow + v];
      sn[v] = SN[crow + v];
    } else {
      own[v] = reg[v * TREG + t];
    }
  }
  const double Vol = M.vol[L];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const double o = rk_update(irk, own[v], sn[v], rhs[v], dt, Vol);
    Sout[crow + v] = o;
    if (SfOut) SfOut[crow + v] = (float)o;
    if (write_rhs) RHS[crow + v] = rhs[v];
  }
}

/* Green-Gauss gradients of the interior tiles' cells: k_grad's viscous rows of the
 * momentum gradients (and, with heat conduction, the gradient of p / rho) from the tile
 * and the cell layer beyond each of its faces, staged once in LDS (no neighbour-index
 * loads, coalesced rows).  The arithmetic is k_grad's, term for term, so the values are
 * bitwise the same.  Halo slot
 * of face f: 512 + f x 64 + u x 8 + w, (u, w) the other two coordinates in axis order. */
template <int MODEL, typename R, bool HEAT, typename SS = double>
/* FP32 state (the mixed form): the rows are staged in FP32 (each value converts exactly
 * to the double the arithmetic uses), 25 KB of LDS and at most 64 VGPRs, so that four
 * workgroups share a CU instead of three */
__global__ __launch_bounds__(TT, sizeof(SS) == 4 ? 4 : 1) void k_tile_grad(DevMesh M, DevParams P,
                                                  const int32_t *__restrict__ tiles,
                                                  const SS *__restrict__ S,
                                                  R *__restrict__ SGm) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  constexpr int NV = HEAT ? NEQ : 3; /* staged rows: all (theta needs them) or the momenta */
  constexpr int V0 = HEAT ? 0 : 1;   /* first staged variable */
  constexpr int HR = TT + 384;
  __shared__ SS reg[NV * HR];
  __shared__ double th[HEAT ? HR : 1];
  const int t = threadIdx.x;
  const int32_t *row = tiles + (size_t)logical_block() * 8;
  const TileLinks tb{row[0], row[1], row[2], row[3], row[4], row[5], row[6]};
  const int L = row[7];
#pragma unroll
  for (int v = 0; v < NV; ++v) reg[v * HR + t] = S[(size_t)(tb.base + t) * NEQ + V0 + v];
  if (t < 384) {
    const int f = t >> 6, u = (t >> 3) & 7, w = t & 7;
    int cell;
    switch (f) {
      case 0: cell = tb.my + tmorton(u, 7, w); break;
      case 1: cell = tb.px + tmorton(0, u, w); break;
      case 2: cell = tb.py + tmorton(u, 0, w); break;
      case 3: cell = tb.mx + tmorton(7, u, w); break;
      case 4: cell = tb.mz + tmorton(u, w, 7); break;
      default: cell = tb.pz + tmorton(u, w, 0); break;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) reg[v * HR + TT + t] = S[(size_t)cell * NEQ + V0 + v];
  }
  __syncthreads();
  if constexpr (HEAT) {
    double U[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; ++v) U[v] = (double)reg[v * HR + t];
    th[t] = theta_ideal(P, U);
    if (t < 384) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) U[v] = (double)reg[v * HR + TT + t];
      th[TT + t] = theta_ideal(P, U);
    }
    __syncthreads();
  }
  int x, y, z;
  demorton(t, x, y, z);
  /* slot of the neighbour across face f */
  auto nslot = [&](int f) -> int {
    switch (f) {
      case 0: return y > 0 ? tmorton(x, y - 1, z) : TT + 0 * 64 + x * 8 + z;
      case 1: return x < 7 ? tmorton(x + 1, y, z) : TT + 1 * 64 + y * 8 + z;
      case 2: return y < 7 ? tmorton(x, y + 1, z) : TT + 2 * 64 + x * 8 + z;
      case 3: return x > 0 ? tmorton(x - 1, y, z) : TT + 3 * 64 + y * 8 + z;
      case 4: return z > 0 ? tmorton(x, y, z - 1) : TT + 4 * 64 + x * 8 + y;
      default: return z < 7 ? tmorton(x, y, z + 1) : TT + 5 * 64 + x * 8 + y;
    }
  };
  const double V = M.vol[L];
  double U[4];
#pragma unroll
  for (int v = 1; v < 4; ++v) U[v] = (double)reg[(v - V0) * HR + t];
  const double town = HEAT ? th[t] : 0.0;
  const int i = tb.base + t;
  double gm[9]; /* gm[3 i + d] = d(rho u_i)/dx_d */
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int fa = d == 0 ? 1 : (d == 1 ? 0 : 4);
    const int fb = d == 0 ? 3 : (d == 1 ? 2 : 5);
    double acc[4], acct = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int f = k == 0 ? fa : fb;
      const double nd = (f == 1 || f == 2 || f == 5) ? 1.0 : -1.0; /* c_face_sign */
      const double A = M.area[3 * L + d];
      const double wgt = nd * A / V;
      const int r = nslot(f);
#pragma unroll
      for (int v = 1; v < 4; ++v) acc[v] += 0.5 * ((double)reg[(v - V0) * HR + r] + U[v]) * wgt;
      if constexpr (HEAT) acct += 0.5 * (th[r] + town) * wgt;
    }
    if constexpr (HEAT) SGm[((size_t)VROW_HEAT + d) * M.ns + i] = (R)acct;
#pragma unroll
    for (int v = 1; v < 4; ++v) gm[(v - 1) * 3 + d] = acc[v];
  }
  double vr[6];
  vrows<double>(gm, vr);
#pragma unroll
  for (int k = 0; k < 6; ++k) SGm[(size_t)k * M.ns + i] = (R)vr[k];
}

template <int MODEL, typename R>
void launch_tile_grad_r(fvg_ctx *c, const double *S, const float *Sf) {
  DevMesh M = dev_mesh(c);
  if constexpr (MODEL == 1 && sizeof(R) == 4) {
    if (Sf) {
      /* the mixed ideal-gas stage: from the FP32 copy of S */
      if (c->dp.kheat != 0.0)
        hipLaunchKernelGGL((k_tile_grad<1, R, true, float>), dim3(c->ntile), dim3(TT), 0, c->stream, M,
                           c->dp, (const int32_t *)c->tiles, Sf, (R *)c->SGm);
      else
        hipLaunchKernelGGL((k_tile_grad<1, R, false, float>), dim3(c->ntile), dim3(TT), 0, c->stream, M,
                           c->dp, (const int32_t *)c->tiles, Sf, (R *)c->SGm);
      return;
    }
  }
  if constexpr (MODEL == 1) {
    if (c->dp.kheat != 0.0) {
      hipLaunchKernelGGL((k_tile_grad<1, R, true>), dim3(c->ntile), dim3(TT), 0, c->stream, M, c->dp,
                         (const int32_t *)c->tiles, S, (R *)c->SGm);
      return;
    }
  }
  hipLaunchKernelGGL((k_tile_grad<MODEL, R, false>), dim3(c->ntile), dim3(TT), 0, c->stream, M, c->dp,
                     (const int32_t *)c->tiles, S, (R *)c->SGm);
}

template <int MODEL, bool NS, typename T, typename R, int OPT, bool FAST>
void launch_tiles_t(fvg_ctx *c, const double *S, const double *SN, double *Sout, float *SfOut,
                    int irk, double dt, bool write_rhs) {
  DevMesh M = dev_mesh(c);
  hipLaunchKernelGGL((k_tile_stage<MODEL, NS, T, R, OPT, FAST>), dim3(c->ntile), dim3(TT), 0,
                     c->stream, M, c->dp, (const int32_t *)c->tiles, S, (const float *)c->Sf, SN,
                     (const R *)c->SGm, irk, dt, Sout, SfOut, c->RHS, write_rhs ? 1 : 0, c->err);
}

template <int MODEL, bool NS>
void launch_tiles_m(fvg_ctx *c, const double *S, const double *SN, double *Sout, float *SfOut,
                    int irk, double dt, bool write_rhs, bool fast) {
  const bool heat = c->dp.kheat != 0.0;
  if (c->rsize == 8) {
    if constexpr (MODEL == 1 && NS) {
      if (heat) return launch_tiles_t<MODEL, NS, double, double, OPT_HEAT, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
    }
    return launch_tiles_t<MODEL, NS, double, double, 0, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
  }
  if constexpr (MODEL == 1) {
    if (fast) {
      if constexpr (NS) {
        if (heat) return launch_tiles_t<1, NS, float, float, OPT_HEAT, true>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
      }
      return launch_tiles_t<1, NS, float, float, 0, true>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
    }
    if constexpr (NS) {
      if (heat) return launch_tiles_t<1, NS, float, float, OPT_HEAT, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
    }
  }
  launch_tiles_t<MODEL, NS, float, float, 0, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs);
}

}  // namespace

bool fv_tiles_usable(const fvg_ctx *c) {
  /* MUSCL (the tile evaluates regular faces only), no option but heat conduction */
  if (c->tiling == FVG_TILING_OFF || c->ntile == 0 || c->prm.ischeme != 2 || c->prm.iprimtv != 0 ||
      c->prm.wave_speeds != FVG_WAVES_DAVIS)
    return false;
  if (c->tiling == FVG_TILING_ON) return true;
  /* Auto, from measurements on 128^3 meshes (step times, tiled / face-list): Taylor-Green
   * (Navier-Stokes, every leaf in a tile) 0.65 plain / 0.72 mixed, Sod (2/3 of the leaves
   * in tiles) 0.91 / 1.01, the barotropic cavitation 1.07 / 1.15; the 64^3 Sod mesh (216
   * tiles, 42% of the leaves) 1.35 / 1.56: a tile is one workgroup's serial work, and the
   * face-list kernels still sweep the blocks around the tiles.  So: the ideal gas, plain
   * or viscous, with at least 256 interior tiles holding half of the leaves. */
  return c->prm.model == 1 && (c->prm.ns || c->rsize == 8) && c->ntile >= 256 &&
         (int64_t)c->ntile * TT * 2 >= c->N;
}

void fv_launch_tile_stage(fvg_ctx *c, const double *S, const double *SN, double *Sout, float *SfOut,
                          int irk, double dt, bool write_rhs, bool fast) {
  switch (c->prm.model) {
    case 1:
      if (c->prm.ns) launch_tiles_m<1, true>(c, S, SN, Sout, SfOut, irk, dt, write_rhs, fast);
      else launch_tiles_m<1, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs, fast);
      break;
    case 2: launch_tiles_m<2, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs, false); break;
    default:
      if (c->prm.ns) launch_tiles_m<3, true>(c, S, SN, Sout, SfOut, irk, dt, write_rhs, false);
      else launch_tiles_m<3, false>(c, S, SN, Sout, SfOut, irk, dt, write_rhs, false);
      break;
  }
}

void fv_launch_tile_grad(fvg_ctx *c, const double *S, const float *Sf) {
  if (c->prm.model == 2) return; /* no viscous form */
  if (c->prm.model == 1) {
    if (c->rsize == 8) launch_tile_grad_r<1, double>(c, S, nullptr);
    else launch_tile_grad_r<1, float>(c, S, Sf);
  } else {
    if (c->rsize == 8) launch_tile_grad_r<3, double>(c, S, nullptr);
    else launch_tile_grad_r<3, float>(c, S, nullptr);
  }
}

void fv_tiles_release(fvg_ctx *c) {
  if (c->tiles) (void)hipFree(c->tiles);
  if (c->face_skip) (void)hipFree(c->face_skip);
  if (c->gather_skip) (void)hipFree(c->gather_skip);
  if (c->tile_cell) (void)hipFree(c->tile_cell);
  c->tiles = nullptr;
  c->face_skip = c->gather_skip = c->tile_cell = nullptr;
  c->ntile = 0;
  c->face_all_skip = c->gather_all_skip = false;
}

/* ------------------------------------------------------------------ host: tiles */

/* The interior tiles of a mesh and the skip flags of the stage's face kernel and cell
 * blocks (host only); cellmask (one byte per cell, 1 in an interior tile) is filled when a
 * cell block is covered in part. */
static void find_tiles(const fvg_mesh *m, std::vector<int32_t> &rows, std::vector<uint8_t> &fskip,
                       std::vector<uint8_t> &gskip, std::vector<uint8_t> &cellmask, bool &fall,
                       bool &gall) {
  rows.clear();
  fskip.clear();
  gskip.clear();
  cellmask.clear();
  fall = gall = false;
  const int N = m->ncell;
  if (N < TT) return;
  auto level = [&](int i) { return (int)((m->cflag[i] >> 16) & 0xff); };
  auto cls = [&](int i, int f) { return (int)((m->cflag[i] >> (2 * f)) & 3); };
  auto nbr = [&](int f, int i) { return m->nbr[(size_t)f * N + i]; };
  /* candidate tiles: 512 consecutive cells of one level whose +x / +y / +z neighbours
   * inside the block are the Morton ones.  Refinement elsewhere shifts the blocks of a
   * region by any count of leaves, so every start is tried; a start fails within its
   * first few cells unless it begins a block. */
  std::vector<int32_t> tile_of((size_t)N, -1);
  std::vector<int32_t> base;
  static const int fplus[3] = {1, 2, 5};
  for (int b = 0; b + TT <= N;) {
    const int L = level(b);
    bool ok = level(b + TT - 1) == L;
    for (int t = 0; t < TT && ok; ++t) {
      const int i = b + t;
      if (level(i) != L) { ok = false; break; }
      int xyz[3];
      demorton(t, xyz[0], xyz[1], xyz[2]);
      for (int a = 0; a < 3 && ok; ++a) {
        if (xyz[a] == 7) continue;
        int q[3] = {xyz[0], xyz[1], xyz[2]};
        ++q[a];
        const int f = fplus[a];
        ok = cls(i, f) == CLS_SAME && nbr(f, i) == b + tmorton(q[0], q[1], q[2]);
      }
    }
    if (!ok) { ++b; continue; }
    for (int t = 0; t < TT; ++t) tile_of[(size_t)b + t] = (int32_t)base.size();
    base.push_back(b);
    b += TT;
  }
  if (base.empty()) return;
  /* interior tiles: across each face the six cell rows continue into another tile, at
   * the Morton positions of its opposite side */
  std::vector<uint8_t> interior((size_t)N, 0);
  for (size_t k = 0; k < base.size(); ++k) {
    const int b = base[k];
    int32_t nb[6];
    bool ok = true;
    for (int f = 0; f < 6 && ok; ++f) {
      const int a = c_face_axis_host(f);
      const bool plus = f == 1 || f == 2 || f == 5;
      int tn = -1;
      for (int u = 0; u < 8 && ok; ++u)
        for (int w = 0; w < 8 && ok; ++w) {
          int in[3], out[3];
          const int e = plus ? 7 : 0, o = plus ? 0 : 7;
          if (a == 0) { in[0] = e; in[1] = u; in[2] = w; out[0] = o; out[1] = u; out[2] = w; }
          else if (a == 1) { in[0] = u; in[1] = e; in[2] = w; out[0] = u; out[1] = o; out[2] = w; }
          else { in[0] = u; in[1] = w; in[2] = e; out[0] = u; out[1] = w; out[2] = o; }
          const int i = b + tmorton(in[0], in[1], in[2]);
          if (cls(i, f) != CLS_SAME) { ok = false; break; }
          const int n = nbr(f, i);
          if (n < 0 || n >= N || tile_of[(size_t)n] < 0) { ok = false; break; }
          const int t2 = tile_of[(size_t)n];
          if (tn < 0) tn = t2;
          if (t2 != tn || n != base[(size_t)t2] + tmorton(out[0], out[1], out[2])) ok = false;
        }
      nb[f] = ok ? base[(size_t)tn] : -1;
    }
    if (!ok) continue;
    rows.push_back(b);
    for (int f = 0; f < 6; ++f) rows.push_back(nb[f]);
    rows.push_back(level(b));
    for (int t = 0; t < TT; ++t) interior[(size_t)b + t] = 1;
  }
  if (rows.empty()) return;
  /* skip flags: stage face-kernel blocks (per axis group, FV_BLOCK faces each) and
   * gather blocks (FV_BLOCK cells) that hold only interior-tile faces / cells */
  fall = gall = true;
  int64_t start = 0;
  for (int a = 0; a < 3; ++a) {
    const int n = m->nface_axis[a];
    for (int q0 = 0; q0 < n; q0 += FV_BLOCK) {
      bool all = true;
      for (int q = q0; q < n && q < q0 + FV_BLOCK && all; ++q) {
        const int64_t j = start + q;
        all = interior[(size_t)m->face_L[j]] && interior[(size_t)m->face_R[j]];
      }
      fskip.push_back(all ? 1 : 0);
      fall = fall && all;
    }
    start += n;
  }
  bool partial = false;
  for (int i0 = 0; i0 < N; i0 += FV_BLOCK) {
    int in = 0, n = 0;
    for (int i = i0; i < N && i < i0 + FV_BLOCK; ++i, ++n) in += interior[(size_t)i];
    const int flag = in == n ? 1 : (in ? 2 : 0);
    gskip.push_back((uint8_t)flag);
    gall = gall && flag == 1;
    partial = partial || flag == 2;
  }
  /* the gather's grid is padded to a multiple of 8 blocks; the padding has no cells */
  while (gskip.size() % 8) gskip.push_back(1);
  if (partial) cellmask.swap(interior);
}

int fv_tiles_build(fvg_ctx *c, const fvg_mesh *m) {
  fv_tiles_release(c);
  std::vector<int32_t> rows;
  std::vector<uint8_t> fskip, gskip, cellmask;
  bool fall, gall;
  find_tiles(m, rows, fskip, gskip, cellmask, fall, gall);
  const int nt = (int)(rows.size() / 8);
  if (nt == 0) return FVG_OK;
  hipError_t e = hipMalloc((void **)&c->tiles, rows.size() * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&c->face_skip, fskip.size());
  if (e == hipSuccess) e = hipMalloc((void **)&c->gather_skip, gskip.size());
  if (e == hipSuccess && !cellmask.empty()) {
    e = hipMalloc((void **)&c->tile_cell, cellmask.size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->tile_cell, cellmask.data(), cellmask.size(), hipMemcpyHostToDevice, c->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->tiles, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->face_skip, fskip.data(), fskip.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->gather_skip, gskip.data(), gskip.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    fv_tiles_release(c);
    return set_err(c, FVG_ERR_HIP, hipGetErrorString(e));
  }
  c->ntile = nt;
  c->face_all_skip = fall;
  c->gather_all_skip = gall;
  return FVG_OK;
}

extern "C" int fvh_tile_table(const fvg_mesh *m, int32_t *rows, int cap, int *ntile) {
  if (!m || !ntile || (cap > 0 && !rows)) return FVG_ERR_ARG;
  std::vector<int32_t> r;
  std::vector<uint8_t> fskip, gskip, cellmask;
  bool fall, gall;
  find_tiles(m, r, fskip, gskip, cellmask, fall, gall);
  *ntile = (int)(r.size() / 8);
  const int n = *ntile < cap ? *ntile : cap;
  if (n > 0) memcpy(rows, r.data(), (size_t)n * 8 * sizeof(int32_t));
  return FVG_OK;
}
