/*
 * fvmesh.hip — the flattened mesh built on the device from the leaf keys of an
 * adapted forest, and the cell state carried over from the previous mesh by key.
 *
 * The host adapts the forest (split / merge / 2:1 balance with the reference's
 * rules) and hands over only each leaf's key, in list order: its level L and its
 * global cell index (i, j, k) at that level (the roots form an nroot[0] x nroot[1]
 * x nroot[2] grid of level-1 cells).  Everything the host flattener derives by
 * walking the BRANCH structs is then found on the device:
 *   k_rm_insert    key -> cell hash table (open addressing, linear probing)
 *   k_rm_classify  per cell and face: the same-level cell across the face (wrapped
 *                  on periodic axes), else the 4 finer cells of the sub-faces, else
 *                  the coarser cell (the lookups of the forest's resolve_face, with
 *                  only leaves in the table); class, neighbour, the count of finer
 *                  faces, and the leaf that first sets each per-level table entry
 *   k_rm_count     per cell and axis: its owned faces of each kind (regular,
 *                  domain boundary, hanging, other same-level)
 *   k_scan_*       exclusive sums of the 13 count columns (quad and face offsets)
 *   k_rm_fill      quads, faces at their final positions (per axis the kinds in
 *                  that order, each in owner order, as the host flattener orders
 *                  them), gather slots, regular stencils
 *   k_rm_carry     state of each new cell from the old mesh: same key -> copy; an
 *                  old ancestor -> copy (split.c: kids copy the parent's S); the 8
 *                  old kids -> their mean summed in kid order (merge.c:101)
 * The arrays equal the host flattener's for the same forest, bit for bit, and so
 * do the tables: the host evaluates them from the leaves named here, exactly as
 * its own scan would.  No geometry is re-validated: the keys come from a forest
 * of this library, whose cells have the level's size by construction.
 */
#include <climits>
#include <cstring>
#include <vector>

#include "fv_common.h"

/* a named namespace: profilers print kernels of an anonymous one without a name */
namespace fvmesh {

constexpr uint64_t KEY_EMPTY = ~0ull;
constexpr int KEY_BITS = 19; /* global cell index per axis < 2^19 */
constexpr int MAX_LEVEL = 30;
constexpr int SCAN_TILE = 2048; /* cells per scan block: 8 per thread */
constexpr int NKIND = 4;        /* face kinds in an axis group: regular, boundary, hanging, other */
constexpr int NCOL = 3 * NKIND + 1; /* count columns: kinds x 3 axes, then finer faces */

enum { RM_DUP = 1, RM_BALANCE = 2, RM_BC = 4, RM_SELF = 8, RM_CARRY = 16 };

__constant__ int c_fnodes[6][4] = {
    {0, 1, 5, 4}, {3, 7, 5, 1}, {2, 6, 7, 3}, {2, 0, 4, 6}, {2, 3, 1, 0}, {6, 4, 5, 7}};
__constant__ int c_fopp[6] = {2, 3, 0, 1, 5, 4};
__constant__ int c_faxis[6] = {1, 0, 1, 0, 2, 2};
__constant__ int c_fsgn[6] = {-1, 1, 1, -1, -1, 1};
__constant__ int c_fside[6] = {2, 1, 3, 0, 4, 5};
/* the faces of each axis in ascending order */
__constant__ int c_axf[3][2] = {{1, 3}, {0, 2}, {4, 5}};

struct Geo {
  int nroot[3], periodic[3], bc[6];
};

__host__ __device__ inline uint64_t pack_key(int L, int i, int j, int k) {
  return ((uint64_t)L << (3 * KEY_BITS)) | ((uint64_t)i << (2 * KEY_BITS)) |
         ((uint64_t)j << KEY_BITS) | (uint64_t)k;
}

__device__ inline void unpack_key(uint64_t key, int &L, int n[3]) {
  const uint64_t m = (1ull << KEY_BITS) - 1;
  L = (int)(key >> (3 * KEY_BITS));
  n[0] = (int)((key >> (2 * KEY_BITS)) & m);
  n[1] = (int)((key >> KEY_BITS) & m);
  n[2] = (int)(key & m);
}

__device__ inline uint32_t hslot(uint64_t k, int bits) {
  return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

/* cell of key k, or -1; the table is at most half full, and the probe is bounded */
__device__ inline int hfind(const uint64_t *__restrict__ hk, const int32_t *__restrict__ hv,
                            int bits, uint64_t k) {
  const uint32_t mask = (1u << bits) - 1;
  uint32_t s = hslot(k, bits);
  for (uint32_t p = 0; p <= mask; ++p) {
    const uint64_t x = hk[s];
    if (x == k) return hv[s];
    if (x == KEY_EMPTY) return -1;
    s = (s + 1) & mask;
  }
  return -1;
}

__device__ inline int kid_bits(int c, int d) { return (c >> d) & 1; }

/* the same-level cell across face f of cell (L, n): false on a non-periodic
 * domain boundary (the forest's `across`) */
__device__ inline bool across(const Geo &g, int L, const int n[3], int f, int m[3]) {
  const int a = c_faxis[f];
  const int Na = g.nroot[a] << (L - 1);
  m[0] = n[0]; m[1] = n[1]; m[2] = n[2];
  m[a] += c_fsgn[f];
  if (m[a] < 0 || m[a] >= Na) {
    if (!g.periodic[a]) return false;
    m[a] = (m[a] + Na) % Na;
  }
  return true;
}

/* level L + 1 still has its cell indices inside the key range */
__device__ inline bool finer_fits(const Geo &g, int L) {
  return L < MAX_LEVEL && (g.nroot[0] << L) <= (1 << KEY_BITS) &&
         (g.nroot[1] << L) <= (1 << KEY_BITS) && (g.nroot[2] << L) <= (1 << KEY_BITS);
}

/* key of kid k (sub-face order) of the same-level cell m across face f */
__device__ inline uint64_t subface_key(int L, const int m[3], int f, int k) {
  const int c = c_fnodes[f][k] ^ (1 << c_faxis[f]);
  return pack_key(L + 1, 2 * m[0] + kid_bits(c, 0), 2 * m[1] + kid_bits(c, 1),
                  2 * m[2] + kid_bits(c, 2));
}

__global__ void k_fill_u64v(uint64_t *p, size_t n, uint64_t v) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    p[q] = v;
}

__global__ void k_fill_i32v(int32_t *p, size_t n, int32_t v) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    p[q] = v;
}

__global__ __launch_bounds__(256) void k_rm_insert(int n, const uint64_t *__restrict__ keys,
                                                   uint64_t *hk, int32_t *hv, int bits,
                                                   int *__restrict__ err) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const uint32_t mask = (1u << bits) - 1;
  uint32_t s = hslot(k, bits);
  for (uint32_t p = 0; p <= mask; ++p) {
    const unsigned long long prev =
        atomicCAS((unsigned long long *)&hk[s], (unsigned long long)KEY_EMPTY, (unsigned long long)k);
    if (prev == KEY_EMPTY) { hv[s] = i; return; }
    if (prev == k) { atomicOr(err, RM_DUP); return; }
    s = (s + 1) & mask;
  }
}

/* class and neighbour of each face of cell i; the cell's owner entries go to sown */
__device__ inline void classify_cell(int i, int n, int ns, const uint64_t *__restrict__ keys,
                                     const uint64_t *__restrict__ hk,
                                     const int32_t *__restrict__ hv, int bits, const Geo &g,
                                     uint32_t *__restrict__ cflag, int32_t *__restrict__ nbr,
                                     int32_t *__restrict__ cnt, int *sown,
                                     int *__restrict__ err) {
  int L, c[3];
  unpack_key(keys[i], L, c);
  uint32_t cf = (uint32_t)L << 16;
  int nfin = 0;
  int e = 0;
  for (int f = 0; f < 6; ++f) {
    int m[3], cls, nb;
    if (!across(g, L, c, f, m)) {
      const int bc = g.bc[c_fside[f]];
      if (bc != 1 && bc != 2) e |= RM_BC;
      cls = CLS_BOUNDARY;
      nb = -1 - (bc == 1 ? 1 : 2);
    } else if ((nb = hfind(hk, hv, bits, pack_key(L, m[0], m[1], m[2]))) >= 0) {
      cls = CLS_SAME;
      if (nb == i) e |= RM_SELF;
    } else if (finer_fits(g, L) && hfind(hk, hv, bits, subface_key(L, m, f, 0)) >= 0) {
      cls = CLS_FINER;
      nb = -1;
      for (int k = 1; k < 4; ++k)
        if (hfind(hk, hv, bits, subface_key(L, m, f, k)) < 0) e |= RM_BALANCE;
      ++nfin;
    } else if (L > 1 && (nb = hfind(hk, hv, bits, pack_key(L - 1, m[0] >> 1, m[1] >> 1, m[2] >> 1))) >= 0) {
      cls = CLS_COARSER;
      if (nb == i) e |= RM_SELF;
    } else {
      e |= RM_BALANCE;
      cls = CLS_BOUNDARY;
      nb = -2;
    }
    cf |= (uint32_t)cls << (2 * f);
    nbr[(size_t)f * ns + i] = nb;
  }
  cflag[i] = cf;
  cnt[(size_t)(3 * NKIND) * n + i] = nfin;
  if (e) atomicOr(err, e);
  atomicMin(&sown[L], i);
  if (L > 1) atomicMin(&sown[L - 1], i);
  atomicMin(&sown[32 + L], i);
  for (int a = 0; a < 3; ++a) {
    const int fa = c_axf[a][0], fb = c_axf[a][1];
    const bool fin_a = ((cf >> (2 * fa)) & 3) == CLS_FINER, fin_b = ((cf >> (2 * fb)) & 3) == CLS_FINER;
    if (!fin_a || !fin_b) atomicMin(&sown[64 + 3 * L + a], i);
    if ((fin_a || fin_b) && L + 1 < 32) atomicMin(&sown[64 + 3 * (L + 1) + a], i);
  }
}

/* owners: [0, 32) h rows (leaves of level L or L + 1), [32, 64) volumes (level L),
 * [64, 160) areas (level lf of the finer side, axis a).  Every cell names itself for up
 * to 8 entries and a whole level shares one entry, so the minima are taken in LDS first:
 * a block sends at most one global atomicMin per entry it touched, and none where the
 * entry already holds an earlier cell (a stale read only costs an extra atomic). */
constexpr int NOWN = 160;
static_assert(NOWN == FVG_REMESH_OWNERS, "owner table size");

__global__ __launch_bounds__(256) void k_rm_classify(int n, int ns, const uint64_t *__restrict__ keys,
                                                     const uint64_t *__restrict__ hk,
                                                     const int32_t *__restrict__ hv, int bits,
                                                     Geo g, uint32_t *__restrict__ cflag,
                                                     int32_t *__restrict__ nbr,
                                                     int32_t *__restrict__ cnt,
                                                     int32_t *__restrict__ own,
                                                     int *__restrict__ err) {
  __shared__ int sown[NOWN];
  for (int t = threadIdx.x; t < NOWN; t += 256) sown[t] = INT_MAX;
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) classify_cell(i, n, ns, keys, hk, hv, bits, g, cflag, nbr, cnt, sown, err);
  __syncthreads();
  for (int t = threadIdx.x; t < NOWN; t += 256) {
    const int v = sown[t];
    if (v != INT_MAX && v < own[t]) atomicMin(&own[t], v);
  }
}

/* a same-level face of cell i (owner) through face f to nb is regular when both
 * cells have same-level neighbours on their far sides along the axis */
__device__ inline bool regular_face(const uint32_t *__restrict__ cflag, uint32_t cf, int f, int nb) {
  return ((cf >> (2 * c_fopp[f])) & 3) == CLS_SAME && ((cflag[nb] >> (2 * f)) & 3) == CLS_SAME;
}

__global__ __launch_bounds__(256) void k_rm_count(int n, int ns, const uint32_t *__restrict__ cflag,
                                                  const int32_t *__restrict__ nbr,
                                                  int32_t *__restrict__ cnt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t cf = cflag[i];
  for (int a = 0; a < 3; ++a) {
    int k[NKIND] = {0, 0, 0, 0};
    for (int t = 0; t < 2; ++t) {
      const int f = c_axf[a][t];
      const int cls = (cf >> (2 * f)) & 3;
      const int nb = nbr[(size_t)f * ns + i];
      if (cls == CLS_BOUNDARY) k[1] += 1;
      else if (cls == CLS_FINER) k[2] += 4;
      else if (cls == CLS_SAME && nb > i) k[regular_face(cflag, cf, f, nb) ? 0 : 3] += 1;
    }
    for (int q = 0; q < NKIND; ++q) cnt[(size_t)(NKIND * a + q) * n + i] = k[q];
  }
}

/* exclusive prefix of v over a 256-thread block; total receives the block's sum */
__device__ inline int block_exclusive_scan(int v, int &total) {
  __shared__ int ws[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < w) pre += ws[q];
    tot += ws[q];
  }
  __syncthreads();
  total = tot;
  return pre + x - v;
}

/* grid (ntile, NCOL): the sum of each tile of each column */
__global__ __launch_bounds__(256) void k_scan_tiles(const int32_t *__restrict__ cnt, int n, int ntile,
                                                    int32_t *__restrict__ tsum) {
  const int tile = blockIdx.x, col = blockIdx.y;
  const size_t base = (size_t)col * n;
  int s = 0;
  for (int q = threadIdx.x; q < SCAN_TILE; q += 256) {
    const int idx = tile * SCAN_TILE + q;
    if (idx < n) s += cnt[base + idx];
  }
  int tot;
  (void)block_exclusive_scan(s, tot);
  if (threadIdx.x == 0) tsum[(size_t)col * ntile + tile] = tot;
}

/* grid (NCOL): tile sums -> exclusive tile offsets, and the column totals */
__global__ __launch_bounds__(256) void k_scan_sums(int32_t *__restrict__ tsum, int ntile,
                                                   int32_t *__restrict__ tot) {
  const int col = blockIdx.x;
  int carry = 0;
  for (int b0 = 0; b0 < ntile; b0 += 256) {
    const int idx = b0 + threadIdx.x;
    const int v = idx < ntile ? tsum[(size_t)col * ntile + idx] : 0;
    int t;
    const int ex = block_exclusive_scan(v, t);
    if (idx < ntile) tsum[(size_t)col * ntile + idx] = carry + ex;
    carry += t;
  }
  if (threadIdx.x == 0) tot[col] = carry;
}

/* grid (ntile, NCOL): exclusive offsets of every cell */
__global__ __launch_bounds__(256) void k_scan_apply(const int32_t *__restrict__ cnt, int n, int ntile,
                                                    const int32_t *__restrict__ tsum,
                                                    int32_t *__restrict__ off) {
  const int tile = blockIdx.x, col = blockIdx.y;
  const size_t base = (size_t)col * n;
  const int i0 = tile * SCAN_TILE + threadIdx.x * 8;
  int v[8], loc = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[q] = i0 + q < n ? cnt[base + i0 + q] : 0;
    loc += v[q];
  }
  int t;
  int ex = block_exclusive_scan(loc, t) + tsum[(size_t)col * ntile + tile];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (i0 + q < n) off[base + i0 + q] = ex;
    ex += v[q];
  }
}

struct FillOut {
  int32_t *nbr, *quad, *fL, *fR, *fLL, *fRR, *gslot;
  uint32_t *fmeta;
  int32_t ks[3 * NKIND]; /* first face of each kind of each axis */
};

__global__ __launch_bounds__(256) void k_rm_fill(int n, int ns, const uint64_t *__restrict__ keys,
                                                 const uint64_t *__restrict__ hk,
                                                 const int32_t *__restrict__ hv, int bits, Geo g,
                                                 const uint32_t *__restrict__ cflag,
                                                 const int32_t *__restrict__ off, FillOut o) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int L, c[3];
  unpack_key(keys[i], L, c);
  const uint32_t cf = cflag[i];
  /* quads, in face order, from this cell's offset */
  int qf[6];
  int q = off[(size_t)(3 * NKIND) * n + i];
  for (int f = 0; f < 6; ++f) {
    qf[f] = -1;
    if (((cf >> (2 * f)) & 3) != CLS_FINER) continue;
    int m[3];
    across(g, L, c, f, m);
    qf[f] = q;
    o.nbr[(size_t)f * ns + i] = q;
    for (int k = 0; k < 4; ++k) o.quad[4 * q + k] = hfind(hk, hv, bits, subface_key(L, m, f, k));
    ++q;
  }
  for (int a = 0; a < 3; ++a) {
    int pos[NKIND];
    for (int k = 0; k < NKIND; ++k) pos[k] = o.ks[NKIND * a + k] + off[(size_t)(NKIND * a + k) * n + i];
    for (int t = 0; t < 2; ++t) {
      const int f = c_axf[a][t], fo = c_fopp[f];
      const int cls = (cf >> (2 * f)) & 3;
      if (cls == CLS_BOUNDARY) {
        const int bc = -1 - o.nbr[(size_t)f * ns + i];
        const int j = pos[1]++;
        o.fL[j] = i;
        o.fR[j] = i;
        o.fmeta[j] = (uint32_t)f | ((uint32_t)f << 3) | ((uint32_t)bc << 6) | ((uint32_t)L << 8);
        o.fLL[j] = -1;
        o.fRR[j] = -1;
        o.gslot[(size_t)f * ns + i] = j << 1;
      } else if (cls == CLS_FINER) {
        o.gslot[(size_t)f * ns + i] = pos[2] << 1;
        for (int k = 0; k < 4; ++k) {
          const int r = o.quad[4 * qf[f] + k];
          const int j = pos[2]++;
          o.fL[j] = i;
          o.fR[j] = r;
          o.fmeta[j] = (uint32_t)f | ((uint32_t)fo << 3) | ((uint32_t)(L + 1) << 8);
          o.fLL[j] = -1;
          o.fRR[j] = -1;
          o.gslot[(size_t)fo * ns + r] = (j << 1) | 1;
        }
      } else if (cls == CLS_SAME) {
        const int nb = o.nbr[(size_t)f * ns + i];
        if (nb <= i) continue;
        const bool reg = regular_face(cflag, cf, f, nb);
        const int j = reg ? pos[0]++ : pos[3]++;
        o.fL[j] = i;
        o.fR[j] = nb;
        o.fmeta[j] = (uint32_t)f | ((uint32_t)fo << 3) | ((uint32_t)L << 8);
        o.fLL[j] = reg ? o.nbr[(size_t)fo * ns + i] : -1;
        o.fRR[j] = reg ? o.nbr[(size_t)f * ns + nb] : -1;
        o.gslot[(size_t)f * ns + i] = j << 1;
        o.gslot[(size_t)fo * ns + nb] = (j << 1) | 1;
      }
    }
  }
}

/* one lane per (cell, variable) of the new mesh */
__global__ __launch_bounds__(256) void k_rm_carry(int n, int neq, Geo g, const uint64_t *__restrict__ keys,
                                                  const uint64_t *__restrict__ ohk,
                                                  const int32_t *__restrict__ ohv, int obits,
                                                  const double *__restrict__ So,
                                                  double *__restrict__ Sn, int *__restrict__ err) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)n * neq) return;
  const int i = (int)(t / neq), v = (int)(t - (size_t)i * neq);
  int L, c[3];
  unpack_key(keys[i], L, c);
  int o = hfind(ohk, ohv, obits, pack_key(L, c[0], c[1], c[2]));
  for (int d = 1; o < 0 && d < L; ++d)
    o = hfind(ohk, ohv, obits, pack_key(L - d, c[0] >> d, c[1] >> d, c[2] >> d));
  if (o >= 0) {
    Sn[t] = So[(size_t)o * neq + v];
    return;
  }
  double acc = 0.0;
  for (int k = 0; k < 8; ++k) {
    const int q = finer_fits(g, L) ? hfind(ohk, ohv, obits,
                                        pack_key(L + 1, 2 * c[0] + kid_bits(k, 0),
                                                 2 * c[1] + kid_bits(k, 1), 2 * c[2] + kid_bits(k, 2)))
                                : -1;
    if (q < 0) {
      atomicOr(err, RM_CARRY);
      return;
    }
    acc += So[(size_t)q * neq + v] / ((double)8);
  }
  Sn[t] = acc;
}

inline int grid_of(size_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace fvmesh
using namespace fvmesh;

struct fv_remesh {
  int n = 0, ns = 0, bits = 0;
  Geo g;
  uint64_t *keys = nullptr, *hk = nullptr;
  int32_t *hv = nullptr;
  uint32_t *cflag = nullptr;
  int32_t *nbr = nullptr, *cnt = nullptr, *off = nullptr, *tsum = nullptr;
  int32_t *info = nullptr; /* err, NCOL totals, 160 owners */
  int32_t tot[NCOL];
};

void fv_remesh_release(fvg_ctx *c) {
  if (!c || !c->rm) return;
  fv_remesh *r = c->rm;
  void *ptrs[] = {r->keys, r->hk, r->hv, r->cflag, r->nbr, r->cnt, r->off, r->tsum, r->info};
  for (void *p : ptrs) fv_dfree(c, p);
  delete r;
  c->rm = nullptr;
}

static const char *remesh_msg(int e) {
  if (e & RM_DUP) return "device re-flatten: two leaves with one key";
  if (e & RM_BALANCE) return "device re-flatten: a face without a leaf within one level (2:1 violated)";
  if (e & RM_BC) return "device re-flatten: domain boundary with bc other than 1 or 2";
  if (e & RM_SELF) return "device re-flatten: a cell is its own neighbour";
  return "device re-flatten: a new cell has no source in the old mesh";
}

extern "C" {

int fvg_remesh_begin(fvg_ctx *c, int32_t n, const int32_t *level, const int32_t *ijk,
                     const int32_t nroot[3], const int32_t periodic[3], const int32_t bc[6],
                     int32_t *owners) {
  if (!c || !level || !ijk || !nroot || !periodic || !bc || !owners) return FVG_ERR_ARG;
  if (n <= 0 || n > (1 << 26)) return set_err(c, FVG_ERR_ARG, "remesh: cell count out of range");
  Geo g;
  for (int a = 0; a < 3; ++a) {
    if (nroot[a] <= 0 || nroot[a] >= (1 << KEY_BITS))
      return set_err(c, FVG_ERR_ARG, "remesh: bad root grid");
    g.nroot[a] = nroot[a];
    g.periodic[a] = periodic[a] ? 1 : 0;
  }
  for (int s = 0; s < 6; ++s) g.bc[s] = bc[s];
  /* keys on the host: level and index range checked here */
  std::vector<uint64_t> keys((size_t)n);
  for (int i = 0; i < n; ++i) {
    const int L = level[i];
    if (L < 1 || L > MAX_LEVEL) return set_err(c, FVG_ERR_ARG, "remesh: leaf level out of range");
    for (int a = 0; a < 3; ++a) {
      const int64_t Na = (int64_t)nroot[a] << (L - 1);
      const int x = ijk[3 * i + a];
      if (Na > (1 << KEY_BITS) || x < 0 || x >= Na)
        return set_err(c, FVG_ERR_ARG, "remesh: leaf index outside the domain or the key range");
    }
    keys[i] = pack_key(L, ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
  }
  (void)hipSetDevice(c->device);
  fv_remesh_release(c);
  fv_remesh *r = new fv_remesh();
  c->rm = r;
  r->n = n;
  r->ns = (int)round_up(n, 64);
  r->g = g;
  int bits = 10;
  while ((1ll << bits) < 2ll * n) ++bits;
  r->bits = bits;
  const size_t cap = (size_t)1 << bits, ns = (size_t)r->ns;
  const int ntile = grid_of((size_t)n, SCAN_TILE);
  hipStream_t st = c->stream;
  HIPCHK(c, fv_dmalloc(c, &r->keys, (size_t)n * 8, true));
  HIPCHK(c, fv_dmalloc(c, &r->hk, cap * 8, true));
  HIPCHK(c, fv_dmalloc(c, &r->hv, cap * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->cflag, ns * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->nbr, 6 * ns * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->cnt, (size_t)NCOL * n * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->off, (size_t)NCOL * n * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->tsum, (size_t)NCOL * ntile * 4, true));
  HIPCHK(c, fv_dmalloc(c, &r->info, (size_t)(1 + NCOL + FVG_REMESH_OWNERS) * 4, true));
  HIPCHK(c, hipMemcpyAsync(r->keys, keys.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fill_u64v, dim3(1024), dim3(256), 0, st, r->hk, cap, KEY_EMPTY);
  hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(256), 0, st, r->info, (size_t)1 + NCOL, 0);
  hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(256), 0, st, r->info + 1 + NCOL,
                     (size_t)FVG_REMESH_OWNERS, INT_MAX);
  int *err = r->info;
  int32_t *tot = r->info + 1, *own = r->info + 1 + NCOL;
  const int nb = grid_of((size_t)n, 256);
  hipLaunchKernelGGL(k_rm_insert, dim3(nb), dim3(256), 0, st, n, r->keys, r->hk, r->hv, bits, err);
  hipLaunchKernelGGL(k_rm_classify, dim3(nb), dim3(256), 0, st, n, r->ns, r->keys, r->hk, r->hv, bits,
                     g, r->cflag, r->nbr, r->cnt, own, err);
  hipLaunchKernelGGL(k_rm_count, dim3(nb), dim3(256), 0, st, n, r->ns, r->cflag, r->nbr, r->cnt);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntile, NCOL), dim3(256), 0, st, r->cnt, n, ntile, r->tsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(NCOL), dim3(256), 0, st, r->tsum, ntile, tot);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntile, NCOL), dim3(256), 0, st, r->cnt, n, ntile, r->tsum,
                     r->off);
  HIPCHK(c, hipGetLastError());
  int32_t host[1 + NCOL + FVG_REMESH_OWNERS];
  HIPCHK(c, hipMemcpyAsync(host, r->info, sizeof(host), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  if (host[0]) {
    const int e = host[0];
    fv_remesh_release(c);
    return set_err(c, (e & (RM_BC | RM_SELF)) ? FVG_ERR_UNSUPPORTED : FVG_ERR_STATE, remesh_msg(e));
  }
  int64_t nface = 0;
  for (int k = 0; k < 3 * NKIND; ++k) nface += host[1 + k];
  if (nface >= (1ll << 30)) {
    fv_remesh_release(c);
    return set_err(c, FVG_ERR_UNSUPPORTED, "remesh: too many faces");
  }
  for (int k = 0; k < NCOL; ++k) r->tot[k] = host[1 + k];
  for (int k = 0; k < FVG_REMESH_OWNERS; ++k) owners[k] = host[1 + NCOL + k] == INT_MAX ? -1 : host[1 + NCOL + k];
  return FVG_OK;
}

int fvg_remesh_end(fvg_ctx *c, int32_t nlevel, const double *h, const double *dist,
                   const double *area, const double *vol, int carry_state) {
  if (!c || !h || !dist || !area || !vol) return FVG_ERR_ARG;
  fv_remesh *r = c->rm;
  if (!r) return set_err(c, FVG_ERR_STATE, "remesh_end without remesh_begin");
  if (nlevel < 2 || nlevel > 32) return set_err(c, FVG_ERR_ARG, "remesh: bad table rows");
  if (carry_state && (!c->has_mesh || !c->keys))
    return set_err(c, FVG_ERR_STATE, "remesh: no device-built mesh to carry the state from");
  (void)hipSetDevice(c->device);
  hipStream_t st = c->stream;
  const int n = r->n;
  int32_t nfa[3], nreg[3], ks[3 * NKIND];
  int F = 0;
  for (int a = 0; a < 3; ++a) {
    nreg[a] = r->tot[NKIND * a];
    nfa[a] = 0;
    for (int k = 0; k < NKIND; ++k) {
      ks[NKIND * a + k] = F + nfa[a];
      nfa[a] += r->tot[NKIND * a + k];
    }
    F += nfa[a];
  }
  const int Q = r->tot[3 * NKIND];
  /* the old state and keys stay alive for the carry-over */
  double *So = c->S;
  uint64_t *okeys = c->keys, *ohk = c->hkeys;
  int32_t *ohv = c->hval;
  const int obits = c->hbits;
  c->S = nullptr;
  c->keys = c->hkeys = nullptr;
  c->hval = nullptr;
  int rc = fv_alloc_mesh(c, n, F, Q, nlevel, nfa, nreg, false);
  if (!rc) {
    c->cflag = r->cflag;
    c->nbr = r->nbr;
    r->cflag = nullptr;
    r->nbr = nullptr;
    FillOut o{c->nbr, c->quad, c->fL, c->fR, c->fLL, c->fRR, c->gslot, c->fmeta, {}};
    memcpy(o.ks, ks, sizeof(ks));
    hipLaunchKernelGGL(k_rm_fill, dim3(grid_of((size_t)n, 256)), dim3(256), 0, st, n, c->ns, r->keys,
                       r->hk, r->hv, r->bits, r->g, c->cflag, r->off, o);
    if (carry_state) {
      hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(64), 0, st, r->info, (size_t)1, 0);
      hipLaunchKernelGGL(k_rm_carry, dim3(grid_of((size_t)n * c->neq, 256)), dim3(256), 0, st, n,
                         c->neq, r->g, r->keys, ohk, ohv, obits, So, c->S, r->info);
    }
    rc = fv_upload_tables(c, nlevel, h, dist, area, vol);
  }
  int carry_err = 0;
  if (!rc && carry_state &&
      hipMemcpyAsync(&carry_err, r->info, 4, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = set_err(c, FVG_ERR_HIP, "remesh: copy failed");
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = set_err(c, FVG_ERR_HIP, "remesh: kernels failed");
  /* the old blocks go back to the cache only once nothing queued can still read them */
  if (rc) (void)hipStreamSynchronize(st);
  void *old[] = {So, okeys, ohk, ohv};
  for (void *p : old) fv_dfree(c, p);
  if (rc) {
    fv_remesh_release(c);
    return rc;
  }
  c->keys = r->keys;
  c->hkeys = r->hk;
  c->hval = r->hv;
  c->hbits = r->bits;
  r->keys = r->hk = nullptr;
  r->hv = nullptr;
  fv_remesh_release(c);
  fv_dcache_trim(c);
  c->has_mesh = true;
  if (carry_err) return set_err(c, FVG_ERR_STATE, remesh_msg(carry_err));
  return FVG_OK;
}

int fvg_has_keys(const fvg_ctx *c) { return c ? (c->keys != nullptr) : FVG_ERR_ARG; }

int fvg_get_mesh(fvg_ctx *c, int32_t *sizes, uint32_t *cflag, int32_t *nbr, int32_t *quad_nbr,
                 int32_t *face_L, int32_t *face_R, uint32_t *face_meta, int32_t *face_LL,
                 int32_t *face_RR, int32_t *gslot, double *tab) {
  if (!c) return FVG_ERR_ARG;
  if (!c->has_mesh) return set_err(c, FVG_ERR_STATE, "no mesh");
  (void)hipSetDevice(c->device);
  hipStream_t st = c->stream;
  const int N = c->N, F = c->F, Q = c->Q;
  if (sizes) {
    const int32_t v[12] = {N, F, c->nfa[0], c->nfa[1], c->nfa[2], Q, c->nlevel,
                           c->nreg[0], c->nreg[1], c->nreg[2], c->keys != nullptr, 0};
    memcpy(sizes, v, sizeof(v));
  }
  const hipMemcpyKind d2h = hipMemcpyDeviceToHost;
  if (cflag) HIPCHK(c, hipMemcpyAsync(cflag, c->cflag, (size_t)N * 4, d2h, st));
  for (int f = 0; f < 6; ++f) {
    if (nbr)
      HIPCHK(c, hipMemcpyAsync(nbr + (size_t)f * N, c->nbr + (size_t)f * c->ns, (size_t)N * 4, d2h, st));
    if (gslot)
      HIPCHK(c, hipMemcpyAsync(gslot + (size_t)f * N, c->gslot + (size_t)f * c->ns, (size_t)N * 4, d2h, st));
  }
  if (quad_nbr && Q) HIPCHK(c, hipMemcpyAsync(quad_nbr, c->quad, (size_t)4 * Q * 4, d2h, st));
  if (face_L) HIPCHK(c, hipMemcpyAsync(face_L, c->fL, (size_t)F * 4, d2h, st));
  if (face_R) HIPCHK(c, hipMemcpyAsync(face_R, c->fR, (size_t)F * 4, d2h, st));
  if (face_meta) HIPCHK(c, hipMemcpyAsync(face_meta, c->fmeta, (size_t)F * 4, d2h, st));
  if (face_LL) HIPCHK(c, hipMemcpyAsync(face_LL, c->fLL, (size_t)F * 4, d2h, st));
  if (face_RR) HIPCHK(c, hipMemcpyAsync(face_RR, c->fRR, (size_t)F * 4, d2h, st));
  if (tab) HIPCHK(c, hipMemcpyAsync(tab, c->tab, (size_t)34 * c->nlevel * 8, d2h, st));
  HIPCHK(c, hipStreamSynchronize(st));
  return FVG_OK;
}

}  // extern "C"
