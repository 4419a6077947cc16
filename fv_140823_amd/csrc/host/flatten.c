/*
 * Flattener: the leaf list of a RUN (reference BRANCH/CELL/LEAF structs) ->
 * the struct-of-arrays mesh the device kernels read (fvg_mesh).
 *
 * Per leaf, in list order (Morton within each root, rebuildlist.c:34-91):
 *   cflag  level and a 2-bit class per face (same level, finer = 4 sub-faces,
 *          coarser, domain boundary);
 *   nbr    the neighbour index per face (quad index for finer faces, -1-bc on
 *          the boundary).
 * Unique faces, grouped by axis (x, y, z); a same-level face is owned by the
 * lower leaf index, a hanging face by the coarse side (4 consecutive sub-faces
 * k = 0..3 in the order of neigtr[f][k], elconnfin.c:41-45), a boundary face by
 * its leaf.  Within an axis the regular faces come first, then the domain boundary
 * faces, the hanging faces and the other same-level faces, each part in owner-leaf
 * order.  A regular face is an interior same-level face whose
 * two cells also have same-level neighbours on their far sides along the axis:
 * its MUSCL stencil is four cells of one level (face_LL, face_L, face_R,
 * face_RR), and the kernels evaluate it without the class and neighbour lookups.
 * gslot[f][i] names the face (and side) whose flux enters leaf i through its
 * face f (for a finer neighbour the first of its 4 consecutive sub-faces, k = 0..3;
 * residual_faces_AMR.c:37-175).
 * Geometry becomes per-level tables: cell size, the distance() factors per
 * neighbour class (distance.c:36-144), face area (meshproperties.c:29-65) and
 * volume (meshproperties.c:66-111).  Leaves that disagree with their level's
 * table by more than 1e-12 relative are rejected (FVG_ERR_UNSUPPORTED).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fvgpu.h"
#include "fv_tables.h"
#include "ptrmap.h"

#define FV_MAXLEV 32

struct fvh_flat {
  fvg_mesh m;
  uint32_t *cflag;
  int32_t *nbr, *quad_nbr, *face_L, *face_R, *face_LL, *face_RR, *gslot;
  uint32_t *face_meta;
  double h[FV_MAXLEV * 3];
  double dist[FV_MAXLEV * 3 * 3 * 2];
  double area[FV_MAXLEV * 3];
  double vol[FV_MAXLEV];
};

/* order of the faces within an axis group: regular, domain boundary, the 4 sub-faces
 * of a hanging face, then the other same-level faces (a stencil across a level
 * change); each kind in owner order.  The kernels' waves then mostly hold faces of
 * one kind and take one branch. */
static int face_kind(int reg, int bc, int owner_cls) {
  return reg ? 0 : (bc != 0 ? 1 : (owner_cls == FV_CLS_FINER ? 2 : 3));
}

static int close_rel(double a, double b) {
  double s = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
  return fabs(a - b) <= 1e-12 * s;
}

/* cell size of branch b along axis a from its nodes (node 0 -> node 1/2/4) */
static double cell_h(const BRANCH *b, int a) {
  const ND *n0 = &b->cl->nd[0];
  const ND *n1 = &b->cl->nd[1 << a];
  return a == 0 ? n1->x - n0->x : (a == 1 ? n1->y - n0->y : n1->z - n0->z);
}

static int set_table(double *slot, double v, unsigned char *have) {
  if (!*have) { *slot = v; *have = 1; return 0; }
  return close_rel(*slot, v) ? 0 : FVG_ERR_UNSUPPORTED;
}

/* Finer levels that only appear as sub-face partners have their h from 2:1 kids; any
 * level without h up to maxlev + 1 takes half the coarser one.  Then the distance.c
 * factors per level L of the cell, axis a, class. */
static void fill_h_and_dist(double *h, unsigned char *hav_h, double *dist, int maxlev) {
  for (int L = 1; L <= maxlev + 1 && L < FV_MAXLEV; ++L)
    for (int a = 0; a < 3; ++a)
      if (!hav_h[3 * L + a] && L > 1 && hav_h[3 * (L - 1) + a]) {
        h[3 * L + a] = 0.5 * h[3 * (L - 1) + a];
        hav_h[3 * L + a] = 1;
      }
  for (int L = 1; L <= maxlev; ++L)
    for (int a = 0; a < 3; ++a) {
      double *d = &dist[((L * 3 + a) * 3) * 2];
      double hL = h[3 * L + a];
      double t = 0.5 * sqrt(hL * hL); /* same level (distance.c:85-87) */
      d[0] = 0.5 * t;
      d[1] = t;
      t = 0.5 * sqrt(hL * hL); /* finer neighbour: its parent's centre (distance.c:112-114) */
      d[2] = 0.5 * t;
      d[3] = 0.5 * t + 0.25 * t;
      double hC = L > 1 ? h[3 * (L - 1) + a] : 0.0;
      t = 0.25 * sqrt(hC * hC); /* coarser neighbour: own parent's centre (distance.c:141-143) */
      d[4] = 0.25 * t;
      d[5] = 0.25 * t + 0.5 * t;
    }
}

int fvh_tables_from_owners(BRANCH *const *lv, int n, const int32_t *own, double *h, double *dist,
                           double *area, double *vol, int32_t *nlevel) {
  if (!lv || n <= 0 || !own || !h || !dist || !area || !vol || !nlevel) return FVG_ERR_ARG;
  memset(h, 0, sizeof(double) * FV_MAXLEV * 3);
  memset(dist, 0, sizeof(double) * FV_MAXLEV * 18);
  memset(area, 0, sizeof(double) * FV_MAXLEV * 3);
  memset(vol, 0, sizeof(double) * FV_MAXLEV);
  unsigned char hav_h[FV_MAXLEV * 3] = {0};
  int maxlev = 0;
  for (int L = 1; L < FV_MAXLEV; ++L) {
    const int iv = own[32 + L], ih = own[L];
    if (iv >= n || ih >= n) return FVG_ERR_ARG;
    if (iv >= 0) {
      if (lv[iv]->level != L) return FVG_ERR_ARG;
      vol[L] = lv[iv]->cl->Vol;
      maxlev = L;
    }
    if (ih >= 0) {
      /* the first leaf of level L or L + 1: its own size, or its parent's */
      const BRANCH *b = lv[ih];
      if (b->level != L && (b->level != L + 1 || !b->prnt)) return FVG_ERR_ARG;
      for (int a = 0; a < 3; ++a) {
        h[3 * L + a] = cell_h(b->level == L ? b : b->prnt, a);
        hav_h[3 * L + a] = 1;
      }
    }
    for (int a = 0; a < 3; ++a) {
      const int ia = own[64 + 3 * L + a];
      if (ia < 0) continue;
      if (ia >= n) return FVG_ERR_ARG;
      /* its first face of axis a whose finer side has level L, first sub-face */
      const BRANCH *b = lv[ia];
      int f = 0;
      for (; f < 6; ++f)
        if (FV_FACE_AXIS[f] == a && (b->nsfc[f] == 4 ? b->level + 1 : b->level) == L) break;
      if (f == 6) return FVG_ERR_ARG;
      area[3 * L + a] = b->cl->Area[f][0];
    }
  }
  if (maxlev == 0 || maxlev >= FV_MAXLEV - 1) return FVG_ERR_ARG;
  fill_h_and_dist(h, hav_h, dist, maxlev);
  *nlevel = maxlev + 2;
  return FVG_OK;
}

void fvh_flat_destroy(fvh_flat *fl) {
  if (!fl) return;
  free(fl->cflag); free(fl->nbr); free(fl->quad_nbr); free(fl->face_L); free(fl->face_R);
  free(fl->gslot); free(fl->face_meta); free(fl->face_LL); free(fl->face_RR);
  free(fl);
}

const fvg_mesh *fvh_flat_mesh(const fvh_flat *fl) { return fl ? &fl->m : NULL; }

int fvh_flatten_run(RUN *run, fvh_flat **out) {
  if (!run || !run->topo || !run->con || !out) return FVG_ERR_ARG;
  const int neq = run->con->neq;
  int N = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt) ++N;
  if (N == 0) return FVG_ERR_ARG;
  fvh_flat *fl = (fvh_flat *)calloc(1, sizeof(*fl));
  if (!fl) return FVG_ERR_NOMEM;
  /* leaf -> list position: the leaves' own structs are only read */
  fv_ptrmap pos = {NULL, NULL, 0};
  BRANCH **lv = (BRANCH **)malloc((size_t)N * sizeof(BRANCH *));
  fl->cflag = (uint32_t *)calloc((size_t)N, sizeof(uint32_t));
  fl->nbr = (int32_t *)malloc((size_t)6 * N * sizeof(int32_t));
  fl->gslot = (int32_t *)malloc((size_t)6 * N * sizeof(int32_t));
  int rc = 0;
  if (!lv || !fl->cflag || !fl->nbr || !fl->gslot || fv_ptrmap_init(&pos, (size_t)N)) {
    rc = FVG_ERR_NOMEM;
    goto fail;
  }
  {
    int i = 0;
    for (BRANCH *b = run->topo->locl; b; b = b->lnxt) {
      fv_ptrmap_put(&pos, b, i);
      lv[i++] = b;
    }
  }
  /* ---- per-level geometry tables */
  unsigned char hav_h[FV_MAXLEV * 3] = {0}, hav_a[FV_MAXLEV * 3] = {0}, hav_v[FV_MAXLEV] = {0};
  int maxlev = 0;
  for (int i = 0; i < N && !rc; ++i) {
    const BRANCH *b = lv[i];
    if (b->type != 0 || b->nlfc != 6 || b->level < 1 || b->level >= FV_MAXLEV - 1) {
      rc = FVG_ERR_UNSUPPORTED; break;
    }
    if (b->level > maxlev) maxlev = b->level;
    for (int a = 0; a < 3 && !rc; ++a) {
      rc = set_table(&fl->h[3 * b->level + a], cell_h(b, a), &hav_h[3 * b->level + a]);
      if (!rc && b->level > 1)
        rc = set_table(&fl->h[3 * (b->level - 1) + a], cell_h(b->prnt, a),
                       &hav_h[3 * (b->level - 1) + a]);
    }
    if (!rc) rc = set_table(&fl->vol[b->level], b->cl->Vol, &hav_v[b->level]);
    for (int f = 0; f < 6 && !rc; ++f) {
      int a = FV_FACE_AXIS[f];
      int lf = b->nsfc[f] == 4 ? b->level + 1 : b->level;
      for (int k = 0; k < b->nsfc[f] && !rc; ++k)
        rc = set_table(&fl->area[3 * lf + a], b->cl->Area[f][k], &hav_a[3 * lf + a]);
    }
  }
  if (rc) goto fail;
  fill_h_and_dist(fl->h, hav_h, fl->dist, maxlev);
  /* ---- per-leaf classes and neighbours */
  int nquad = 0;
  for (int i = 0; i < N && !rc; ++i) {
    BRANCH *b = lv[i];
    uint32_t cf = (uint32_t)b->level << 16;
    for (int f = 0; f < 6; ++f) {
      int bc = b->cl->fc[f].bc;
      int cls;
      int32_t nb;
      BRANCH *n0 = b->neigtr[f][0];
      if (bc != 0 || n0 == NULL || n0->level == 0) {
        if (bc != 1 && bc != 2) { rc = FVG_ERR_UNSUPPORTED; break; }
        cls = FV_CLS_BOUNDARY;
        nb = -1 - bc;
      } else if (b->lfc_neigs[f] == 4) {
        cls = FV_CLS_FINER;
        nb = nquad++;
      } else if (b->lfc_neigs[f] == 1 && n0->level == b->level) {
        cls = FV_CLS_SAME;
        nb = fv_ptrmap_get(&pos, n0);
      } else if (b->lfc_neigs[f] == 1 && n0->level == b->level - 1) {
        cls = FV_CLS_COARSER;
        nb = fv_ptrmap_get(&pos, n0);
      } else {
        rc = FVG_ERR_UNSUPPORTED;
        break;
      }
      if (cls != FV_CLS_BOUNDARY && cls != FV_CLS_FINER && (nb < 0 || nb >= N || nb == i)) {
        rc = FVG_ERR_UNSUPPORTED;
        break;
      }
      cf |= (uint32_t)cls << (2 * f);
      fl->nbr[(size_t)f * N + i] = nb;
    }
    fl->cflag[i] = cf;
  }
  if (rc) goto fail;
  fl->quad_nbr = (int32_t *)malloc((size_t)4 * (nquad ? nquad : 1) * sizeof(int32_t));
  if (!fl->quad_nbr) { rc = FVG_ERR_NOMEM; goto fail; }
  for (int i = 0; i < N && !rc; ++i) {
    BRANCH *b = lv[i];
    for (int f = 0; f < 6; ++f) {
      if (((fl->cflag[i] >> (2 * f)) & 3) != FV_CLS_FINER) continue;
      int q = fl->nbr[(size_t)f * N + i];
      for (int k = 0; k < 4; ++k) {
        BRANCH *kb = b->neigtr[f][k];
        if (!kb || kb->level != b->level + 1) { rc = FVG_ERR_UNSUPPORTED; break; }
        const int32_t kq = fv_ptrmap_get(&pos, kb);
        if (kq < 0) { rc = FVG_ERR_UNSUPPORTED; break; }
        fl->quad_nbr[4 * q + k] = kq;
      }
    }
  }
  if (rc) goto fail;
  /* ---- faces: count, then fill by axis */
  int64_t nfa[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i)
    for (int f = 0; f < 6; ++f) {
      int cls = (fl->cflag[i] >> (2 * f)) & 3;
      int nb = fl->nbr[(size_t)f * N + i];
      int a = FV_FACE_AXIS[f];
      if (cls == FV_CLS_BOUNDARY) nfa[a] += 1;
      else if (cls == FV_CLS_FINER) nfa[a] += 4;
      else if (cls == FV_CLS_SAME && nb > i) nfa[a] += 1;
    }
  int64_t nface64 = nfa[0] + nfa[1] + nfa[2];
  if (nface64 >= (1LL << 30)) { rc = FVG_ERR_UNSUPPORTED; goto fail; }
  int nface = (int)nface64;
  fl->face_L = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
  fl->face_R = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
  fl->face_meta = (uint32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(uint32_t));
  if (!fl->face_L || !fl->face_R || !fl->face_meta) { rc = FVG_ERR_NOMEM; goto fail; }
  for (size_t q = 0; q < (size_t)6 * N; ++q) fl->gslot[q] = -2;
  int j = 0;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < N; ++i)
      for (int f = 0; f < 6; ++f) {
        if (FV_FACE_AXIS[f] != a) continue;
        int cls = (fl->cflag[i] >> (2 * f)) & 3;
        int nb = fl->nbr[(size_t)f * N + i];
        uint32_t lev = (fl->cflag[i] >> 16) & 0xff;
        int fo = FV_FACE_OPP[f];
        if (cls == FV_CLS_BOUNDARY) {
          int bc = -1 - nb;
          fl->face_L[j] = i;
          fl->face_R[j] = i;
          fl->face_meta[j] = (uint32_t)f | ((uint32_t)f << 3) | ((uint32_t)bc << 6) | (lev << 8);
          fl->gslot[(size_t)f * N + i] = j << 1;
          ++j;
        } else if (cls == FV_CLS_FINER) {
          fl->gslot[(size_t)f * N + i] = j << 1;
          for (int k = 0; k < 4; ++k) {
            int r = fl->quad_nbr[4 * nb + k];
            fl->face_L[j] = i;
            fl->face_R[j] = r;
            fl->face_meta[j] = (uint32_t)f | ((uint32_t)fo << 3) | ((lev + 1) << 8);
            fl->gslot[(size_t)fo * N + r] = (j << 1) | 1;
            ++j;
          }
        } else if (cls == FV_CLS_SAME && nb > i) {
          fl->face_L[j] = i;
          fl->face_R[j] = nb;
          fl->face_meta[j] = (uint32_t)f | ((uint32_t)fo << 3) | (lev << 8);
          fl->gslot[(size_t)f * N + i] = j << 1;
          fl->gslot[(size_t)fo * N + nb] = (j << 1) | 1;
          ++j;
        }
      }
  /* every face slot must be covered exactly as its class says */
  for (int i = 0; i < N && !rc; ++i)
    for (int f = 0; f < 6; ++f) {
      int32_t g = fl->gslot[(size_t)f * N + i];
      if (g == -2) { rc = FVG_ERR_UNSUPPORTED; break; }
      int cls = (fl->cflag[i] >> (2 * f)) & 3;
      int jf = g >> 1;
      int other = g & 1;
      if (cls == FV_CLS_COARSER && !other) { rc = FVG_ERR_UNSUPPORTED; break; }
      if (other && fl->face_R[jf] != i) { rc = FVG_ERR_UNSUPPORTED; break; }
      if (!other && fl->face_L[jf] != i) { rc = FVG_ERR_UNSUPPORTED; break; }
    }
  if (rc) goto fail;
  /* ---- regular faces first within each axis (stable), and their outer cells */
  {
    int32_t *perm = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
    int32_t *nL = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
    int32_t *nR = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
    uint32_t *nM = (uint32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(uint32_t));
    fl->face_LL = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
    fl->face_RR = (int32_t *)malloc((size_t)(nface ? nface : 1) * sizeof(int32_t));
    if (!perm || !nL || !nR || !nM || !fl->face_LL || !fl->face_RR) {
      free(perm); free(nL); free(nR); free(nM);
      rc = FVG_ERR_NOMEM;
      goto fail;
    }
    int64_t start = 0;
    for (int a = 0; a < 3; ++a) {
      const int64_t end = start + nfa[a];
      int64_t nreg = 0;
      int64_t pos = start;
      for (int pass = 0; pass < 4; ++pass) {
        for (int64_t q = start; q < end; ++q) {
          const int iL = fl->face_L[q], iR = fl->face_R[q];
          const uint32_t mt = fl->face_meta[q];
          const int f = mt & 7, fo = (mt >> 3) & 7, bc = (mt >> 6) & 3;
          const uint32_t cL = fl->cflag[iL], cR = fl->cflag[iR];
          const int fL2 = FV_FACE_OPP[f], fR2 = FV_FACE_OPP[fo];
          const int reg = bc == 0 && ((cL >> (2 * f)) & 3) == FV_CLS_SAME &&
                          ((cL >> (2 * fL2)) & 3) == FV_CLS_SAME &&
                          ((cR >> (2 * fR2)) & 3) == FV_CLS_SAME;
          if (face_kind(reg, bc, (cL >> (2 * f)) & 3) != pass) continue;
          perm[q] = (int32_t)pos;
          nL[pos] = iL;
          nR[pos] = iR;
          nM[pos] = mt;
          fl->face_LL[pos] = reg ? fl->nbr[(size_t)fL2 * N + iL] : -1;
          fl->face_RR[pos] = reg ? fl->nbr[(size_t)fR2 * N + iR] : -1;
          ++pos;
        }
        if (pass == 0) nreg = pos - start;
      }
      fl->m.nface_reg[a] = (int32_t)nreg;
      start = end;
    }
    for (size_t q = 0; q < (size_t)6 * N; ++q)
      if (fl->gslot[q] >= 0) fl->gslot[q] = (perm[fl->gslot[q] >> 1] << 1) | (fl->gslot[q] & 1);
    memcpy(fl->face_L, nL, (size_t)nface * sizeof(int32_t));
    memcpy(fl->face_R, nR, (size_t)nface * sizeof(int32_t));
    memcpy(fl->face_meta, nM, (size_t)nface * sizeof(uint32_t));
    free(perm); free(nL); free(nR); free(nM);
  }
  free(lv);
  lv = NULL;
  fv_ptrmap_free(&pos);
  fl->m.ncell = N;
  fl->m.nface = nface;
  fl->m.nface_axis[0] = (int32_t)nfa[0];
  fl->m.nface_axis[1] = (int32_t)nfa[1];
  fl->m.nface_axis[2] = (int32_t)nfa[2];
  fl->m.nquad = nquad;
  fl->m.nlevel = maxlev + 2;
  fl->m.neq = neq;
  fl->m.cflag = fl->cflag;
  fl->m.nbr = fl->nbr;
  fl->m.quad_nbr = fl->quad_nbr;
  fl->m.face_L = fl->face_L;
  fl->m.face_R = fl->face_R;
  fl->m.face_meta = fl->face_meta;
  fl->m.face_LL = fl->face_LL;
  fl->m.face_RR = fl->face_RR;
  fl->m.gslot = fl->gslot;
  fl->m.h = fl->h;
  fl->m.dist = fl->dist;
  fl->m.area = fl->area;
  fl->m.vol = fl->vol;
  *out = fl;
  return FVG_OK;
fail:
  free(lv);
  fv_ptrmap_free(&pos);
  fvh_flat_destroy(fl);
  return rc;
}
