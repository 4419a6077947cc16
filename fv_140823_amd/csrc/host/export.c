/*
 * Tecplot FEPOINT export of the leaves of a RUN, with the node sharing, numbering,
 * averaging and text layout of exportfield_linked_V3.c (one rank).
 *
 * Mesh nodes ("links") are built by visiting the leaves in list order, their faces
 * 0..5 and the four nodes of each face (exportfield_linked_V3.c:213-305): across an
 * interior face whose neighbour is of the same level, the face node is joined with
 * the neighbour's node at the same point; across a face with finer neighbours, face
 * node k is joined with the corner node of the k-th finer neighbour; a face on the
 * domain boundary or with a coarser neighbour joins nothing and only opens a link for
 * a node that has none.  Joining two existing links appends the second link's
 * members to the first and empties the second; links are then numbered in creation
 * order, skipping the emptied ones (:307-331).  A node's values are the means over
 * its members, accumulated in member order as v / count (:382-405): the conserved
 * state, the pressure, and the magnitudes of the gradients of rho, rho u, rho v and
 * of p.  The neighbour's node at the same point is found by exact coordinate
 * equality; across a periodic side (an extension) the points differ, so the face
 * joins nothing there.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "fvgpu.h"
#include "fv_tables.h"
#include "ptrmap.h"

/* Tecplot BRICK node order of the reference's hexahedron nodes (TPNDS[.][0], :83-90) */
static const int TP_BRICK[8] = {0, 1, 3, 2, 4, 5, 7, 6};

typedef struct {
  int32_t head, tail, count;
} link_t;

static int same_point(const ND *a, const ND *b) { return a->x == b->x && a->y == b->y && a->z == b->z; }

static int node_at(const BRANCH *b2, const ND *p) {
  for (int n = 0; n < 8; ++n)
    if (same_point(&b2->cl->nd[n], p)) return n;
  return -1;
}

int fvh_export_tecplot(RUN *run, const char *path, int model, const double *p) {
  if (!run || !run->topo || !run->con || !path || !p) return FVG_ERR_ARG;
  if (model == 0) model = run->con->model;
  const int neq = run->con->neq;
  if (model < 1 || model > 3) return FVG_ERR_UNSUPPORTED;
  int N = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt) ++N;
  if (N == 0) return FVG_ERR_ARG;
  fv_ptrmap pos = {NULL, NULL, 0};  /* leaf -> list position (the leaves are only read) */
  BRANCH **lv = (BRANCH **)malloc((size_t)N * sizeof(BRANCH *));
  int32_t *elnd = (int32_t *)malloc((size_t)8 * N * sizeof(int32_t));   /* (elem,node) -> link */
  int32_t *next = (int32_t *)malloc((size_t)8 * N * sizeof(int32_t));   /* member list links */
  link_t *lk = (link_t *)malloc((size_t)8 * N * sizeof(link_t));
  int32_t *num = (int32_t *)malloc((size_t)8 * N * sizeof(int32_t));    /* link -> node number */
  FILE *fp = NULL;
  int rc = 0;
  if (!lv || !elnd || !next || !lk || !num || fv_ptrmap_init(&pos, (size_t)N)) {
    rc = FVG_ERR_NOMEM;
    goto done;
  }
  {
    int i = 0;
    for (BRANCH *b = run->topo->locl; b; b = b->lnxt) {
      fv_ptrmap_put(&pos, b, i);
      lv[i++] = b;
    }
  }
  for (size_t q = 0; q < (size_t)8 * N; ++q) elnd[q] = -1;
  int nlink = 0;
#define OPEN(slot) do { int32_t s_ = (slot); elnd[s_] = nlink; next[s_] = -1;                 \
                        lk[nlink].head = lk[nlink].tail = s_; lk[nlink].count = 1; ++nlink; } while (0)
#define APPEND(l, slot) do { int32_t s_ = (slot); elnd[s_] = (l); next[s_] = -1;              \
                             next[lk[l].tail] = s_; lk[l].tail = s_; lk[l].count++; } while (0)
  for (int e = 0; e < N; ++e) {
    const BRANCH *b = lv[e];
    for (int f = 0; f < 6; ++f) {
      const BRANCH *n0 = b->neigtr[f][0];
      const int join = b->cl->fc[f].bc == 0 && n0 && n0->level >= b->level;
      const int finer = n0 && n0->level > b->level;
      for (int k = 0; k < 4; ++k) {
        const int node1 = FV_FACE_NODES[f][k];
        const int32_t s1 = 8 * e + node1;
        const BRANCH *b2 = join ? b->neigtr[f][finer ? k : 0] : NULL;
        if (b2 == b) b2 = NULL;  /* a periodic axis one cell long */
        const int32_t e2 = b2 ? fv_ptrmap_get(&pos, b2) : -1;
        const int node2 = e2 >= 0 ? node_at(b2, &b->cl->nd[node1]) : -1;
        if (node2 < 0) {
          if (elnd[s1] == -1) OPEN(s1);
          continue;
        }
        const int32_t s2 = 8 * e2 + node2;
        const int32_t l1 = elnd[s1], l2 = elnd[s2];
        if (l1 == -1 && l2 == -1) {
          OPEN(s1);
          APPEND(elnd[s1], s2);
        } else if (l1 != -1 && l2 == -1) {
          APPEND(l1, s2);
        } else if (l1 == -1 && l2 != -1) {
          APPEND(l2, s1);
        } else if (l1 != l2) {
          for (int32_t s = lk[l2].head; s != -1; s = next[s]) elnd[s] = l1;
          next[lk[l1].tail] = lk[l2].head;
          lk[l1].tail = lk[l2].tail;
          lk[l1].count += lk[l2].count;
          lk[l2].count = 0;
          lk[l2].head = lk[l2].tail = -1;
        }
      }
    }
  }
#undef OPEN
#undef APPEND
  int nnode = 0;
  for (int l = 0; l < nlink; ++l) num[l] = lk[l].count > 0 ? nnode++ : -1;
  for (size_t q = 0; q < (size_t)8 * N; ++q)
    if (elnd[q] < 0) { rc = FVG_ERR_STATE; goto done; }  /* every element node has a link */

  fp = fopen(path, "w");
  if (!fp) { rc = FVG_ERR_IO; goto done; }
  /* preamble and zone (exportfield_linked_V3.c:347-361) */
  if (model == 1)
    fprintf(fp, " VARIABLES =  \"X\" \"Y\" \"Z\" \"iel\" \"part\" \"rho\" \"U\" \"V\" \"W\" \"E\" \"P\" \"G_rho\" \"G_u\" \"G_v\" \"G_p\" ");
  else if (model == 2)
    fprintf(fp, " VARIABLES =  \"X\" \"Y\" \"Z\" \"iel\" \"part\" \"rho\" \"U\" \"V\" \"W\" \"P\" \"G_rho\" \"G_u\" \"G_v\" \"G_p\" ");
  else
    fprintf(fp, " VARIABLES =  \"X\" \"Y\" \"Z\" \"iel\" \"part\" \"rho\" \"U\" \"V\" \"W\" \"rhoY\" \"P\" \"G_rho\" \"G_u\" \"G_v\" \"G_p\" ");
  fprintf(fp, " \n ");
  fprintf(fp, "ZONE T=\"%d\",N=%d,E=%d,F=FEPOINT,ET=BRICK,DT=(double,double,double)\n", 1, nnode, N);
  /* nodes (:377-467) */
  double var[8 + 1 + 4];
  for (int l = 0; l < nlink; ++l) {
    if (lk[l].count <= 0) continue;
    const int cnt = lk[l].count;
    for (int v = 0; v < neq + 5; ++v) var[v] = 0.0;
    for (int32_t s = lk[l].head; s != -1; s = next[s]) {
      const BRANCH *b = lv[s >> 3];
      double *const *SG = b->el->SG;
      for (int v = 0; v < neq; ++v) var[v] += b->el->S[v] / cnt;
      var[neq] += p[s >> 3] / cnt;
      const double g0 = sqrt(pow(SG[0][0], 2.0) + pow(SG[0][1], 2.0) + pow(SG[0][2], 2.0));
      const double g1 = sqrt(pow(SG[1][0], 2.0) + pow(SG[1][1], 2.0) + pow(SG[1][2], 2.0));
      const double g2 = sqrt(pow(SG[2][0], 2.0) + pow(SG[2][1], 2.0) + pow(SG[2][2], 2.0));
      const double gp = sqrt(pow(SG[neq][0], 2.0) + pow(SG[neq][1], 2.0) + pow(SG[neq][2], 2.0));
      var[neq + 1] += g0 / cnt;
      var[neq + 2] += g1 / cnt;
      var[neq + 3] += g2 / cnt;
      var[neq + 4] += gp / cnt;
    }
    const int32_t s0 = lk[l].head;
    const BRANCH *b0 = lv[s0 >> 3];
    const ND *nd = &b0->cl->nd[s0 & 7];
    fprintf(fp, "%e %e %e %d %d ", nd->x, nd->y, nd->z, b0->root, b0->part);
    for (int v = 0; v < neq + 5; ++v) fprintf(fp, " %e ", var[v]);
    fprintf(fp, " \n");
  }
  /* elements (:469-533) */
  for (int e = 0; e < N; ++e) {
    for (int k = 0; k < 8; ++k) fprintf(fp, "%d ", 1 + num[elnd[8 * e + TP_BRICK[k]]]);
    fprintf(fp, "\n");
  }
  if (ferror(fp)) rc = FVG_ERR_IO;
done:
  if (fp && fclose(fp) != 0) rc = FVG_ERR_IO;
  free(lv); free(elnd); free(next); free(lk); free(num);
  fv_ptrmap_free(&pos);
  return rc;
}
