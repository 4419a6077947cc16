/*
 * orc_runcopy.c — reference-shaped copies of a RUN (TEST INFRASTRUCTURE).
 *
 * The library's forests carve their nodes from large blocks with room for fields of
 * their own.  A RUN built by the reference's host code instead holds every struct in an
 * allocation of exactly its own size (createbranch.c:29: malloc(sizeof(BRANCH)), and so
 * on).  This file makes such a RUN from any RUN: every BRANCH (the leaves, their
 * ancestors and the outside node glob), CELL, LEAF, FOREST, CON, PAR, the RUN itself and
 * every array they point to is a malloc of exactly the reference's size followed by
 * guard bytes, and the RUN's fields other than con / par / topo hold garbage, as a RUN
 * on the reference's stack (main.c:22-30) does.  Tests hand the copy to the library
 * (flattener, export, GPU binding) and then check that no guard byte changed, i.e. that
 * nothing was written past the end of a reference-sized struct.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fvgpu.h"

#define ORC_GUARD 64
#define ORC_CANARY 0xA5

typedef struct orc_copy {
  RUN *run;
  void **blk;     /* every allocation ... */
  size_t *len;    /* ... and its size before the guard */
  int n, cap;
  BRANCH **src, **dst;  /* node map (open addressing on the source address) */
  size_t mask;
  int neq;
} orc_copy;

static void *gmalloc(orc_copy *c, size_t bytes) {
  if (c->n == c->cap) {
    int nc = c->cap ? 2 * c->cap : 1024;
    void **b = (void **)realloc(c->blk, (size_t)nc * sizeof(void *));
    size_t *l = b ? (size_t *)realloc(c->len, (size_t)nc * sizeof(size_t)) : NULL;
    if (b) c->blk = b;
    if (!b || !l) return NULL;
    c->len = l;
    c->cap = nc;
  }
  unsigned char *p = (unsigned char *)malloc(bytes + ORC_GUARD);
  if (!p) return NULL;
  memset(p + bytes, ORC_CANARY, ORC_GUARD);
  c->blk[c->n] = p;
  c->len[c->n] = bytes;
  c->n++;
  return p;
}

static size_t slot_of(const orc_copy *c, const BRANCH *b) {
  uint64_t h = (uint64_t)(uintptr_t)b;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  return (size_t)h & c->mask;
}

static BRANCH *map_get(const orc_copy *c, const BRANCH *b) {
  if (!b) return NULL;
  size_t s = slot_of(c, b);
  while (c->src[s]) {
    if (c->src[s] == b) return c->dst[s];
    s = (s + 1) & c->mask;
  }
  return NULL;
}

static int map_put(orc_copy *c, BRANCH *b, BRANCH *d) {
  size_t s = slot_of(c, b);
  while (c->src[s] && c->src[s] != b) s = (s + 1) & c->mask;
  int fresh = c->src[s] == NULL;
  c->src[s] = b;
  c->dst[s] = d;
  return fresh;
}

static double *copy_doubles(orc_copy *c, const double *s, size_t n) {
  double *d = (double *)gmalloc(c, n * sizeof(double));
  if (d && s) memcpy(d, s, n * sizeof(double));
  return d;
}

static int *copy_ints(orc_copy *c, const int *s, size_t n) {
  int *d = (int *)gmalloc(c, n * sizeof(int));
  if (d && s) memcpy(d, s, n * sizeof(int));
  return d;
}

/* the struct fields of one node; pointers to other nodes are fixed up afterwards */
static BRANCH *copy_node(orc_copy *c, const BRANCH *s, int is_leaf) {
  BRANCH *d = (BRANCH *)gmalloc(c, sizeof(BRANCH));
  if (!d) return NULL;
  memcpy(d, s, sizeof(BRANCH));
  d->adrs = NULL; d->neignd = NULL; d->neigag = NULL; d->ipartbound = NULL;
  d->keen = NULL; d->keenfc = NULL; d->keenfcangle = NULL; d->el = NULL; d->cl = NULL;
  d->neigtr = NULL; d->neigfc = NULL; d->lfc_neigs = NULL; d->nsfc = NULL; d->kids = NULL;
  if (s->cl) {
    const CELL *sc = s->cl;
    CELL *dc = (CELL *)gmalloc(c, sizeof(CELL));
    if (!dc) return NULL;
    memcpy(dc, sc, sizeof(CELL));
    dc->nx = copy_doubles(c, sc->nx, 6); dc->ny = copy_doubles(c, sc->ny, 6);
    dc->nz = copy_doubles(c, sc->nz, 6); dc->nxt1 = copy_doubles(c, sc->nxt1, 6);
    dc->nyt1 = copy_doubles(c, sc->nyt1, 6); dc->nzt1 = copy_doubles(c, sc->nzt1, 6);
    dc->nxt2 = copy_doubles(c, sc->nxt2, 6); dc->nyt2 = copy_doubles(c, sc->nyt2, 6);
    dc->nzt2 = copy_doubles(c, sc->nzt2, 6);
    dc->Area = NULL;
    if (sc->Area) {
      dc->Area = (double **)gmalloc(c, 6 * sizeof(double *));
      if (!dc->Area) return NULL;
      for (int f = 0; f < 6; ++f) dc->Area[f] = copy_doubles(c, sc->Area[f], 4);
    }
    dc->nd = NULL;
    if (sc->nd) {
      dc->nd = (ND *)gmalloc(c, 8 * sizeof(ND));
      if (!dc->nd) return NULL;
      memcpy(dc->nd, sc->nd, 8 * sizeof(ND));
    }
    dc->fc = NULL;
    if (sc->fc) {
      dc->fc = (FC *)gmalloc(c, 6 * sizeof(FC));
      if (!dc->fc) return NULL;
      memcpy(dc->fc, sc->fc, 6 * sizeof(FC));
    }
    d->cl = dc;
  }
  if (s->kids && s->nkids > 0) {
    d->kids = (BRANCH **)gmalloc(c, 8 * sizeof(BRANCH *));
    if (!d->kids) return NULL;
    memcpy(d->kids, s->kids, 8 * sizeof(BRANCH *));
  }
  if (is_leaf) {
    const int neq = c->neq;
    d->neigtr = (BRANCH ***)gmalloc(c, 6 * sizeof(BRANCH **));
    if (!d->neigtr) return NULL;
    for (int f = 0; f < 6; ++f) {
      d->neigtr[f] = (BRANCH **)gmalloc(c, 4 * sizeof(BRANCH *));
      if (!d->neigtr[f]) return NULL;
      memcpy(d->neigtr[f], s->neigtr[f], 4 * sizeof(BRANCH *));
    }
    d->neigfc = copy_ints(c, s->neigfc, 6);
    d->lfc_neigs = copy_ints(c, s->lfc_neigs, 6);
    d->nsfc = copy_ints(c, s->nsfc, 6);
    const LEAF *se = s->el;
    LEAF *de = (LEAF *)gmalloc(c, sizeof(LEAF));
    if (!de) return NULL;
    memcpy(de, se, sizeof(LEAF));
    de->S = copy_doubles(c, se->S, (size_t)neq);
    de->SN = copy_doubles(c, se->SN, (size_t)neq);
    de->RHS = copy_doubles(c, se->RHS, (size_t)neq);
    de->SF = (double **)gmalloc(c, 6 * sizeof(double *));
    de->SG = (double **)gmalloc(c, (size_t)(neq + 1) * sizeof(double *));
    if (!de->SF || !de->SG) return NULL;
    for (int f = 0; f < 6; ++f) de->SF[f] = copy_doubles(c, se->SF[f], (size_t)neq);
    for (int v = 0; v <= neq; ++v) de->SG[v] = copy_doubles(c, se->SG[v], 3);
    de->flux_face = NULL;
    de->flux_flag = copy_ints(c, se->flux_flag, 6);
    d->el = de;
  }
  return d;
}

static void release(orc_copy *c) {
  for (int i = 0; i < c->n; ++i) free(c->blk[i]);
  free(c->blk);
  free(c->len);
  free(c->src);
  free(c->dst);
  free(c);
}

/* A reference-shaped deep copy of src (see the file comment); *handle owns it. */
int orc_copy_run(RUN *src, void **handle, RUN **out) {
  if (!src || !src->topo || !src->con || !src->par || !handle || !out) return -1;
  orc_copy *c = (orc_copy *)calloc(1, sizeof(*c));
  if (!c) return -6;
  c->neq = src->con->neq;
  /* every node the copy needs: the leaves, their ancestors, glob */
  int nleaf = 0;
  for (BRANCH *b = src->topo->locl; b; b = b->lnxt) ++nleaf;
  size_t cap = 64;
  while (cap < (size_t)4 * (nleaf + 1)) cap <<= 1;
  c->mask = cap - 1;
  c->src = (BRANCH **)calloc(cap, sizeof(BRANCH *));
  c->dst = (BRANCH **)calloc(cap, sizeof(BRANCH *));
  BRANCH **order = (BRANCH **)malloc(cap * sizeof(BRANCH *));
  int nn = 0, rc = 0;
  if (!c->src || !c->dst || !order) { rc = -6; goto fail; }
  BRANCH *glob = src->topo->glob;
  if (glob) { map_put(c, glob, NULL); order[nn++] = glob; }
  for (BRANCH *b = src->topo->locl; b; b = b->lnxt)
    for (BRANCH *a = b; a && a != glob; a = a->prnt) {
      if (!map_put(c, a, NULL)) break; /* ancestors already listed */
      order[nn++] = a;
      if (a->prnt == a) break;
    }
  for (int i = 0; i < nn; ++i) {
    BRANCH *s = order[i];
    BRANCH *d = copy_node(c, s, s != glob && s->nkids == 0);
    if (!d) { rc = -6; goto fail; }
    map_put(c, s, d);
  }
  /* node pointers */
  for (int i = 0; i < nn; ++i) {
    BRANCH *s = order[i], *d = map_get(c, s);
    d->prnt = s->prnt == s ? d : map_get(c, s->prnt);
    d->next = map_get(c, s->next);
    d->lnxt = map_get(c, s->lnxt);
    d->lprv = map_get(c, s->lprv);
    if (d->kids)
      for (int k = 0; k < 8; ++k) d->kids[k] = map_get(c, s->kids[k]);
    if (d->neigtr)
      for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 4; ++k) d->neigtr[f][k] = map_get(c, s->neigtr[f][k]);
  }
  {
    FOREST *ft = (FOREST *)gmalloc(c, sizeof(FOREST));
    CON *cn = (CON *)gmalloc(c, sizeof(CON));
    PAR *pr = (PAR *)gmalloc(c, sizeof(PAR));
    RUN *run = (RUN *)gmalloc(c, sizeof(RUN));
    if (!ft || !cn || !pr || !run) { rc = -6; goto fail; }
    memset(ft, 0, sizeof(FOREST));
    ft->ntrees = src->topo->ntrees;
    ft->nleaves = src->topo->nleaves;
    ft->tleaves = src->topo->tleaves;
    ft->locl = map_get(c, src->topo->locl);
    ft->glob = map_get(c, glob);
    memcpy(cn, src->con, sizeof(CON));
    memcpy(pr, src->par, sizeof(PAR));
    memset(run, 0xCD, sizeof(RUN)); /* an uninitialised RUN on the reference's stack */
    run->con = cn;
    run->par = pr;
    run->topo = ft;
    c->run = run;
  }
  free(order);
  *handle = c;
  *out = c->run;
  return 0;
fail:
  free(order);
  release(c);
  return rc;
}

/* The number of allocations of the copy whose guard bytes changed. */
int orc_copy_guards(void *handle) {
  orc_copy *c = (orc_copy *)handle;
  if (!c) return -1;
  int bad = 0;
  for (int i = 0; i < c->n; ++i) {
    const unsigned char *g = (const unsigned char *)c->blk[i] + c->len[i];
    for (int k = 0; k < ORC_GUARD; ++k)
      if (g[k] != ORC_CANARY) { ++bad; break; }
  }
  return bad;
}

void orc_copy_free(void *handle) {
  if (handle) release((orc_copy *)handle);
}

/* Leaf fields of any RUN to / from var-major host arrays (layouts as FVG_FIELD_*). */
int orc_run_field(RUN *run, int which, double *buf, int to_leaves) {
  if (!run || !run->topo || !run->con || !buf) return -1;
  const int neq = run->con->neq;
  int N = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt) ++N;
  int i = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt, ++i) {
    LEAF *el = b->el;
    if (which <= FVG_FIELD_RHS) {
      double *p = which == FVG_FIELD_S ? el->S : (which == FVG_FIELD_SN ? el->SN : el->RHS);
      for (int v = 0; v < neq; ++v) {
        if (to_leaves) p[v] = buf[(size_t)v * N + i];
        else buf[(size_t)v * N + i] = p[v];
      }
    } else if (which == FVG_FIELD_SF) {
      for (int f = 0; f < 6; ++f)
        for (int v = 0; v < neq; ++v) {
          const size_t o = ((size_t)f * neq + v) * N + i;
          if (to_leaves) el->SF[f][v] = buf[o];
          else buf[o] = el->SF[f][v];
        }
    } else {
      return -1;
    }
  }
  return 0;
}
