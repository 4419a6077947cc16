/*
 * Host forest of octrees built directly as the reference's BRANCH/CELL/LEAF
 * structures (strdata.h:76-168), with
 *   - neighbour pointers with the semantics of elconn/elconnfin
 *     (elconn.c:25-180, elconnfin.c:19-123): lfc_neigs 1 for a same-level or
 *     coarser neighbour, 4 for finer ones in sub-face order, glob on the domain
 *     boundary (lfc_neigs 1 at root level, 0 below);
 *   - geometry by the formulas of normalvector.c:54-59, tagvector.c:43-53 and
 *     meshproperties.c:15-111 (areas of hanging sub-faces from the finer
 *     neighbour's face, volume as six tetrahedra, centre = node mean);
 *   - the leaf list in depth-first kid order 0..7 (rebuildlist.c:34-91), i.e.
 *     Morton order within each root;
 *   - split by injection of the parent state (split.c:492-502) and merge by the
 *     arithmetic mean of the kids (merge.c:92-107), with 2:1 face balance.
 * Periodic pairing of opposite sides is an extension (the reference has none).
 * Adaptation is incremental (a (level, i, j, k) hash kept current, a worklist for the
 * 2:1 balance, only changed leaves refreshed); its per-leaf passes run on a pool of
 * host threads, and the branch blocks are carved from large mappings and reused.
 */
#define _POSIX_C_SOURCE 200809L
#define _DEFAULT_SOURCE
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include "fvgpu.h"
#include "fv_tables.h"

/* ------------------------------------------------------------------ threads */

/* [0, n) split into chunks over up to FVH_THREADS host threads; fn(ctx, lo, hi) may touch
 * only what belongs to its range, and the result never depends on the split.  Plain
 * pthreads, as the process may already hold another OpenMP runtime (torch's).  The
 * workers are started on first use and kept (starting and joining 15 threads costs about
 * half a millisecond, and an adaptation makes about ten passes); one pass runs at a time,
 * and fn must not itself call par_for. */
#define FVH_THREADS 16

static struct {
  pthread_mutex_t mu;
  pthread_cond_t go, idle;
  int started;            /* workers running */
  unsigned long gen;      /* pass number */
  void (*fn)(void *, int, int);
  void *ctx;
  int n, nchunk;
  int next;               /* next chunk to hand out (atomic) */
  int active;             /* threads inside the current pass (under mu) */
  int left;               /* chunks of the current pass not yet done (under mu) */
} pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER,
          0, 0, NULL, NULL, 0, 0, 0, 0, 0};
static pthread_mutex_t pool_user = PTHREAD_MUTEX_INITIALIZER;

static int pool_chunks(void (*fn)(void *, int, int), void *ctx, int n, int nchunk) {
  int done = 0;
  for (;;) {
    const int k = __atomic_fetch_add(&pool.next, 1, __ATOMIC_RELAXED);
    if (k >= nchunk) return done;
    fn(ctx, (int)((int64_t)n * k / nchunk), (int)((int64_t)n * (k + 1) / nchunk));
    ++done;
  }
}

static void *pool_worker(void *unused) {
  (void)unused;
  unsigned long seen = 0;
  pthread_mutex_lock(&pool.mu);
  for (;;) {
    while (pool.gen == seen) pthread_cond_wait(&pool.go, &pool.mu);
    seen = pool.gen;
    void (*fn)(void *, int, int) = pool.fn;
    void *ctx = pool.ctx;
    const int n = pool.n, nchunk = pool.nchunk;
    ++pool.active;
    pthread_mutex_unlock(&pool.mu);
    const int done = pool_chunks(fn, ctx, n, nchunk);
    pthread_mutex_lock(&pool.mu);
    --pool.active;
    pool.left -= done;
    pthread_cond_broadcast(&pool.idle);
  }
  return NULL;
}

/* A forked child has none of the parent's workers, and the pool's condition variables
 * still count the parent's waiters: they start afresh. */
static void pool_forked(void) {
  pthread_mutex_init(&pool.mu, NULL);
  pthread_mutex_init(&pool_user, NULL);
  pthread_cond_init(&pool.go, NULL);
  pthread_cond_init(&pool.idle, NULL);
  pool.started = 0;
  pool.active = 0;
  pool.left = 0;
}

static void pool_start(int want) {
  static int atfork = 0;
  if (!atfork) {
    atfork = 1;
    (void)pthread_atfork(NULL, NULL, pool_forked);
  }
  while (pool.started < want) {
    pthread_t t;
    pthread_attr_t at;
    if (pthread_attr_init(&at)) return;
    (void)pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
    const int rc = pthread_create(&t, &at, pool_worker, NULL);
    pthread_attr_destroy(&at);
    if (rc) return; /* fewer workers: the caller takes the chunks nobody takes */
    ++pool.started;
  }
}

/* the CPUs this process may run on (its affinity mask, which cgroup CPU sets narrow),
 * at most FVH_THREADS */
static int host_threads(void) {
  static int n = 0;
  if (!n) {
    cpu_set_t cs;
    long c = sched_getaffinity(0, sizeof(cs), &cs) == 0 ? CPU_COUNT(&cs) : sysconf(_SC_NPROCESSORS_ONLN);
    n = c < 1 ? 1 : (c > FVH_THREADS ? FVH_THREADS : (int)c);
  }
  return n;
}

/* grain: the fewest items worth a thread of its own */
static void par_for(int n, int grain, void (*fn)(void *, int, int), void *ctx) {
  int T = host_threads();
  if (grain < 1) grain = 1;
  if (T > n / grain) T = n / grain;
  if (T <= 1) {
    if (n > 0) fn(ctx, 0, n);
    return;
  }
  /* a few chunks per thread, handed out in turn, even out uneven items */
  int nchunk = n / grain;
  if (nchunk > 4 * T) nchunk = 4 * T;
  pthread_mutex_lock(&pool_user);
  pthread_mutex_lock(&pool.mu);
  pool_start(T - 1);
  while (pool.active > 0) pthread_cond_wait(&pool.idle, &pool.mu);
  pool.fn = fn;
  pool.ctx = ctx;
  pool.n = n;
  pool.nchunk = nchunk;
  __atomic_store_n(&pool.next, 0, __ATOMIC_RELAXED);
  pool.left = nchunk;
  ++pool.gen;
  ++pool.active;
  pthread_cond_broadcast(&pool.go);
  pthread_mutex_unlock(&pool.mu);
  const int done = pool_chunks(fn, ctx, n, nchunk);
  pthread_mutex_lock(&pool.mu);
  --pool.active;
  pool.left -= done;
  while (pool.left > 0) pthread_cond_wait(&pool.idle, &pool.mu);
  pthread_mutex_unlock(&pool.mu);
  pthread_mutex_unlock(&pool_user);
}

struct fvh_forest {
  RUN run;
  CON con;
  PAR par;
  FOREST topo;
  BRANCH *glob;
  BRANCH **roots;
  int nroots;
  int nroot[3];
  double x0[3], x1[3];   /* the box */
  int periodic[3];
  int bc[6];
  int neq;
  /* hash (level,i,j,k) -> branch */
  BRANCH **ht;
  uint64_t *hk;
  size_t hcap, hcount;
  /* leaves in list order */
  BRANCH **leaves;
  int nleaves;
  /* branch blocks: carved from large mappings (huge pages where the kernel gives them:
   * a split's new blocks then cost few page faults), freed ones reused */
  size_t bbytes;
  char **slabs;
  int nslab, cslab;
  size_t slab_used;
  void *freelist;
};

/* ------------------------------------------------------------------ allocation */

/* A node of this forest: the reference's BRANCH, then what the forest keeps per node and
 * the reference's struct has no field for: the node's integer address at its level (root
 * grid index scaled by 2^(level-1)) and its position in the leaf list.  Only nodes made
 * by this file are fvh_nodes; the flattener and the binding, which take any RUN, never
 * use these fields. */
typedef struct fvh_node {
  BRANCH b;
  int ijk[3];
  int index;
} fvh_node;

#define IJK(p) (((fvh_node *)(p))->ijk)
#define IDX(p) (((fvh_node *)(p))->index)

static size_t align8(size_t x) { return (x + 7) & ~(size_t)7; }

/* One block per branch: BRANCH + neighbour tables + CELL + LEAF arrays, laid out over
 * blk (zeroed, branch_bytes(neq) long); with blk null only the size is returned. */
static BRANCH *branch_layout(char *blk, int neq, size_t *bytes) {
  size_t sz = 0;
  size_t o_br = sz; sz = align8(sz + sizeof(fvh_node));
  size_t o_ntr = sz; sz = align8(sz + 6 * sizeof(BRANCH **));
  size_t o_ntr_rows = sz; sz = align8(sz + 6 * 4 * sizeof(BRANCH *));
  size_t o_ints = sz; sz = align8(sz + 4 * 6 * sizeof(int));
  size_t o_kids = sz; sz = align8(sz + 8 * sizeof(BRANCH *));
  size_t o_cl = sz; sz = align8(sz + sizeof(CELL));
  size_t o_norm = sz; sz = align8(sz + 9 * 6 * sizeof(double));
  size_t o_area = sz; sz = align8(sz + 6 * sizeof(double *));
  size_t o_area_rows = sz; sz = align8(sz + 6 * 4 * sizeof(double));
  size_t o_nd = sz; sz = align8(sz + 8 * sizeof(ND));
  size_t o_fc = sz; sz = align8(sz + 6 * sizeof(FC));
  size_t o_el = sz; sz = align8(sz + sizeof(LEAF));
  size_t o_s = sz; sz = align8(sz + 3 * (size_t)neq * sizeof(double));
  size_t o_sf = sz; sz = align8(sz + 6 * sizeof(double *));
  size_t o_sf_rows = sz; sz = align8(sz + 6 * (size_t)neq * sizeof(double));
  size_t o_sg = sz; sz = align8(sz + (size_t)(neq + 1) * sizeof(double *));
  size_t o_sg_rows = sz; sz = align8(sz + (size_t)(neq + 1) * 3 * sizeof(double));
  size_t o_ff = sz; sz = align8(sz + 6 * sizeof(int));
  sz = (sz + 63) & ~(size_t)63;
  if (bytes) *bytes = sz;
  if (!blk) return NULL;
  BRANCH *b = (BRANCH *)(blk + o_br);
  b->neigtr = (BRANCH ***)(blk + o_ntr);
  for (int f = 0; f < 6; ++f) b->neigtr[f] = (BRANCH **)(blk + o_ntr_rows) + 4 * f;
  int *ints = (int *)(blk + o_ints);
  b->neigfc = ints;
  b->lfc_neigs = ints + 6;
  b->nsfc = ints + 12;
  b->neigag = ints + 18;
  b->kids = (BRANCH **)(blk + o_kids);
  CELL *cl = (CELL *)(blk + o_cl);
  double *nv = (double *)(blk + o_norm);
  cl->nx = nv; cl->ny = nv + 6; cl->nz = nv + 12;
  cl->nxt1 = nv + 18; cl->nyt1 = nv + 24; cl->nzt1 = nv + 30;
  cl->nxt2 = nv + 36; cl->nyt2 = nv + 42; cl->nzt2 = nv + 48;
  cl->Area = (double **)(blk + o_area);
  for (int f = 0; f < 6; ++f) cl->Area[f] = (double *)(blk + o_area_rows) + 4 * f;
  cl->nd = (ND *)(blk + o_nd);
  cl->fc = (FC *)(blk + o_fc);
  b->cl = cl;
  LEAF *el = (LEAF *)(blk + o_el);
  double *s = (double *)(blk + o_s);
  el->S = s; el->SN = s + neq; el->RHS = s + 2 * neq;
  el->SF = (double **)(blk + o_sf);
  for (int f = 0; f < 6; ++f) el->SF[f] = (double *)(blk + o_sf_rows) + (size_t)neq * f;
  el->SG = (double **)(blk + o_sg);
  for (int v = 0; v <= neq; ++v) el->SG[v] = (double *)(blk + o_sg_rows) + 3 * v;
  el->flux_flag = (int *)(blk + o_ff);
  el->flux_face = NULL;
  b->el = el;
  b->type = 0;
  b->nlnd = 8;
  b->nlfc = 6;
  b->nsplit = 0;
  IDX(b) = -1;
  return b;
}

#define FVH_SLAB ((size_t)64 << 20)

/* a block for one branch, not yet zeroed (sequential callers only) */
static char *block_take(struct fvh_forest *fr) {
  if (fr->freelist) {
    char *b = (char *)fr->freelist;
    fr->freelist = *(void **)b;
    return b;
  }
  if (!fr->nslab || fr->slab_used + fr->bbytes > FVH_SLAB) {
    if (fr->nslab == fr->cslab) {
      int c = fr->cslab ? 2 * fr->cslab : 8;
      char **q = (char **)realloc(fr->slabs, (size_t)c * sizeof(char *));
      if (!q) return NULL;
      fr->slabs = q;
      fr->cslab = c;
    }
    void *m = mmap(NULL, FVH_SLAB, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (m == MAP_FAILED) return NULL;
#ifdef MADV_HUGEPAGE
    (void)madvise(m, FVH_SLAB, MADV_HUGEPAGE);
#endif
    fr->slabs[fr->nslab++] = (char *)m;
    fr->slab_used = 0;
  }
  char *b = fr->slabs[fr->nslab - 1] + fr->slab_used;
  fr->slab_used += fr->bbytes;
  return b;
}

static void block_give(struct fvh_forest *fr, void *b) {
  *(void **)b = fr->freelist;
  fr->freelist = b;
}

/* a zeroed branch over block blk (touches only blk: may run on any thread) */
static BRANCH *branch_make(const struct fvh_forest *fr, char *blk) {
  memset(blk, 0, fr->bbytes);
  return branch_layout(blk, fr->neq, NULL);
}

/* ------------------------------------------------------------------ hash */

static uint64_t hkey(int level, const int ijk[3]) {
  return ((uint64_t)(level + 1) << 57) | ((uint64_t)(uint32_t)ijk[0] << 38) |
         ((uint64_t)(uint32_t)ijk[1] << 19) | (uint64_t)(uint32_t)ijk[2];
}
static uint64_t hmix(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
  return k;
}
static int ht_reserve(struct fvh_forest *fr, size_t need) {
  if (fr->hcap >= 2 * need + 16) return 0;
  size_t cap = 1024;
  while (cap < 2 * need + 16) cap <<= 1;
  BRANCH **ht = (BRANCH **)calloc(cap, sizeof(BRANCH *));
  uint64_t *hk = (uint64_t *)calloc(cap, sizeof(uint64_t));
  if (!ht || !hk) { free(ht); free(hk); return FVG_ERR_NOMEM; }
  for (size_t i = 0; i < fr->hcap; ++i) {
    if (!fr->ht[i]) continue;
    size_t p = hmix(fr->hk[i]) & (cap - 1);
    while (ht[p]) p = (p + 1) & (cap - 1);
    ht[p] = fr->ht[i]; hk[p] = fr->hk[i];
  }
  free(fr->ht); free(fr->hk);
  fr->ht = ht; fr->hk = hk; fr->hcap = cap;
  return 0;
}
static int ht_insert(struct fvh_forest *fr, BRANCH *b) {
  if (fr->hcap < 2 * (fr->hcount + 1) + 16) {
    int rc = ht_reserve(fr, fr->hcount + 1);
    if (rc) return rc;
  }
  uint64_t k = hkey(b->level, IJK(b));
  size_t p = hmix(k) & (fr->hcap - 1);
  while (fr->ht[p]) {
    if (fr->hk[p] == k) { fr->ht[p] = b; return 0; }
    p = (p + 1) & (fr->hcap - 1);
  }
  fr->ht[p] = b; fr->hk[p] = k; fr->hcount++;
  return 0;
}
/* removal with backward shift: an entry after the hole moves into it unless its
 * home slot lies cyclically between the hole and itself, so probe chains stay whole */
static void ht_remove(struct fvh_forest *fr, const BRANCH *b) {
  if (!fr->hcap) return;
  const size_t mask = fr->hcap - 1;
  uint64_t k = hkey(b->level, IJK(b));
  size_t p = hmix(k) & mask;
  while (fr->ht[p] && fr->hk[p] != k) p = (p + 1) & mask;
  if (!fr->ht[p]) return;
  for (size_t q = (p + 1) & mask; fr->ht[q]; q = (q + 1) & mask) {
    size_t h = hmix(fr->hk[q]) & mask;
    if (((q - h) & mask) >= ((q - p) & mask)) {
      fr->ht[p] = fr->ht[q];
      fr->hk[p] = fr->hk[q];
      p = q;
    }
  }
  fr->ht[p] = NULL;
  fr->hcount--;
}
static BRANCH *ht_find(const struct fvh_forest *fr, int level, const int ijk[3]) {
  if (!fr->hcap) return NULL;
  uint64_t k = hkey(level, ijk);
  size_t p = hmix(k) & (fr->hcap - 1);
  while (fr->ht[p]) {
    if (fr->hk[p] == k) return fr->ht[p];
    p = (p + 1) & (fr->hcap - 1);
  }
  return NULL;
}

/* ------------------------------------------------------------------ geometry */

static int ncell_axis(const struct fvh_forest *fr, int level, int a) {
  return fr->nroot[a] << (level - 1);
}

static void set_nodes(const struct fvh_forest *fr, BRANCH *b) {
  double L[3];
  int N[3];
  for (int a = 0; a < 3; ++a) {
    L[a] = fr->x1[a] - fr->x0[a];
    N[a] = ncell_axis(fr, b->level, a);
  }
  for (int n = 0; n < 8; ++n) {
    int bit[3] = {n & 1, (n >> 1) & 1, (n >> 2) & 1};
    double x[3];
    for (int a = 0; a < 3; ++a)
      x[a] = fr->x0[a] + ((double)(IJK(b)[a] + bit[a]) * L[a]) / (double)N[a];
    b->cl->nd[n].num = n;
    b->cl->nd[n].x = x[0];
    b->cl->nd[n].y = x[1];
    b->cl->nd[n].z = x[2];
  }
}

/* normalvector.c:54-59 and tagvector.c:43-53 */
static void set_normals(BRANCH *b) {
  CELL *cl = b->cl;
  for (int f = 0; f < 6; ++f) {
    const ND *p0 = &cl->nd[FV_FACE_NODES[f][0]];
    const ND *p1 = &cl->nd[FV_FACE_NODES[f][1]];
    const ND *p2 = &cl->nd[FV_FACE_NODES[f][2]];
    double cx = (p1->y - p0->y) * (p2->z - p0->z) - (p2->y - p0->y) * (p1->z - p0->z);
    double cy = (p1->x - p0->x) * (p2->z - p0->z) - (p2->x - p0->x) * (p1->z - p0->z);
    double cz = (p1->x - p0->x) * (p2->y - p0->y) - (p2->x - p0->x) * (p1->y - p0->y);
    double norm = sqrt(cx * cx + cy * cy + cz * cz);
    cl->nx[f] = cx / norm;
    cl->ny[f] = -cy / norm;
    cl->nz[f] = cz / norm;
    double ex = p1->x - p0->x, ey = p1->y - p0->y, ez = p1->z - p0->z;
    double n1 = sqrt(ey * ey + ex * ex + ez * ez);
    cl->nxt1[f] = ex / n1;
    cl->nyt1[f] = ey / n1;
    cl->nzt1[f] = ez / n1;
    double qx = cl->ny[f] * cl->nzt1[f] - cl->nyt1[f] * cl->nz[f];
    double qy = cl->nx[f] * cl->nzt1[f] - cl->nxt1[f] * cl->nz[f];
    double qz = cl->nx[f] * cl->nyt1[f] - cl->nxt1[f] * cl->ny[f];
    double n2 = sqrt(qx * qx + qy * qy + qz * qz);
    cl->nxt2[f] = qx / n2;
    cl->nyt2[f] = -qy / n2;
    cl->nzt2[f] = qz / n2;
  }
}

/* area of face fopp of `o` as two triangles (meshproperties.c:54-62) */
static double face_area(const BRANCH *o, int fopp) {
  double A = 0.0;
  const ND *nd = o->cl->nd;
  for (int it = 0; it < 2; ++it) {
    const ND *q0 = &nd[FV_FACE_NODES[fopp][0]];
    const ND *q1 = &nd[FV_FACE_NODES[fopp][1 + it]];
    const ND *q2 = &nd[FV_FACE_NODES[fopp][2 + it]];
    double dx1 = q1->x - q0->x, dx2 = q2->x - q0->x;
    double dy1 = q1->y - q0->y, dy2 = q2->y - q0->y;
    double dz1 = q1->z - q0->z, dz2 = q2->z - q0->z;
    double a = dy1 * dz2 - dz1 * dy2, bb = dz1 * dx2 - dx1 * dz2, c = dx1 * dy2 - dy1 * dx2;
    A += fabs(0.5 * sqrt(a * a + bb * bb + c * c));
  }
  return A;
}

/* volume as six tetrahedra and centre as the node mean (meshproperties.c:64-111);
 * depends on the nodes only */
static void set_volume(BRANCH *b) {
  CELL *cl = b->cl;
  double V = 0.0;
  for (int it = 0; it < 6; ++it) {
    double dx[3], dy[3], dz[3];
    for (int in = 0; in < 3; ++in) {
      const ND *q = &cl->nd[FV_VOL_TETS[it][in]];
      dx[in] = q->x - cl->nd[0].x;
      dy[in] = q->y - cl->nd[0].y;
      dz[in] = q->z - cl->nd[0].z;
    }
    double c0 = dy[0] * dz[1] - dz[0] * dy[1];
    double c1 = dz[0] * dx[1] - dx[0] * dz[1];
    double c2 = dx[0] * dy[1] - dy[0] * dx[1];
    V += (c0 * dx[2] + c1 * dy[2] + c2 * dz[2]) / 6.0;
  }
  cl->Vol = V;
  cl->xc = cl->yc = cl->zc = 0.0;
  for (int n = 0; n < 8; ++n) {
    cl->xc += cl->nd[n].x / 8;
    cl->yc += cl->nd[n].y / 8;
    cl->zc += cl->nd[n].z / 8;
  }
}

/* meshproperties.c:15-111: face areas (a hanging sub-face from the finer
 * neighbour's face), then volume and centre */
static void set_props(BRANCH *b) {
  CELL *cl = b->cl;
  for (int f = 0; f < 6; ++f) {
    for (int k = 0; k < 4; ++k) cl->Area[f][k] = 0.0;
    for (int k = 0; k < b->nsfc[f]; ++k) {
      if (b->nsfc[f] == 1) cl->Area[f][k] = face_area(b, f);
      else cl->Area[f][k] = face_area(b->neigtr[f][k], b->neigfc[f]);
    }
  }
  set_volume(b);
}

/* ------------------------------------------------------------------ topology */

static int64_t pow10_i64(int e) {
  int64_t r = 1;
  while (e-- > 0) r *= 10;
  return r;
}

static BRANCH *new_branch(const struct fvh_forest *fr, char *blk, BRANCH *prnt, int level,
                          const int ijk[3], int root) {
  BRANCH *b = branch_make(fr, blk);
  b->level = level;
  b->prnt = prnt;
  b->root = root;
  for (int a = 0; a < 3; ++a) IJK(b)[a] = ijk[a];
  for (int f = 0; f < 6; ++f) {
    b->cl->fc[f].num = f;
    b->cl->fc[f].type = 0;
    b->cl->fc[f].nnds = 4;
    b->cl->fc[f].bc = 0;
  }
  set_nodes(fr, b);
  return b;
}

/* The 8 kids of leaf b over the blocks blk[0..8), initialised but not yet in the tree
 * or the hash (b->kids is filled, b->nkids stays 0).  Touches only b and the blocks, so
 * different leaves may be done on different threads. */
static void split_kids(const struct fvh_forest *fr, BRANCH *b, char *const *blk) {
  for (int c = 0; c < 8; ++c) {
    int ijk[3] = {2 * IJK(b)[0] + (c & 1), 2 * IJK(b)[1] + ((c >> 1) & 1),
                  2 * IJK(b)[2] + ((c >> 2) & 1)};
    BRANCH *k = new_branch(fr, blk[c], b, b->level + 1, ijk, b->root);
    for (int v = 0; v < fr->neq; ++v) {
      k->el->S[v] = b->el->S[v];
      k->el->SG[v][0] = b->el->SG[v][0];
      k->el->SG[v][1] = b->el->SG[v][1];
      k->el->SG[v][2] = b->el->SG[v][2];
    }
    /* tagaddress.c:19-27: decimal digit level-2 of the tag is 1 + the kid index at
       that level (split.c:73); past 10 levels it no longer fits an int */
    int64_t t = b->tag < 0 ? -1 : (int64_t)b->tag + pow10_i64(b->level - 1) * (1 + c);
    k->tag = (t >= 0 && t <= INT32_MAX) ? (int)t : -1;
    b->kids[c] = k;
  }
}

/* 8 blocks for one split; none taken on failure */
static int take8(struct fvh_forest *fr, char **blk) {
  for (int c = 0; c < 8; ++c)
    if (!(blk[c] = block_take(fr))) {
      while (c-- > 0) block_give(fr, blk[c]);
      return FVG_ERR_NOMEM;
    }
  return 0;
}

/* kids made by split_kids into the hash and the tree */
static int split_link(struct fvh_forest *fr, BRANCH *b) {
  for (int c = 0; c < 8; ++c) {
    int rc = ht_insert(fr, b->kids[c]);
    if (rc) {
      for (int q = 0; q < c; ++q) ht_remove(fr, b->kids[q]);
      for (int q = 0; q < 8; ++q) {
        block_give(fr, b->kids[q]);
        b->kids[q] = NULL;
      }
      return rc;
    }
  }
  b->nkids = 8;
  b->nsplit = 8;
  return 0;
}

/* a leaf's kids may already have been made (split_kids) and wait to be linked */
static int split_branch(struct fvh_forest *fr, BRANCH *b) {
  if (b->nkids) return 0;
  if (!b->kids[0]) {
    char *blk[8];
    int rc = take8(fr, blk);
    if (rc) return rc;
    split_kids(fr, b, blk);
  }
  return split_link(fr, b);
}

static void free_subtree(struct fvh_forest *fr, BRANCH *b) {
  for (int c = 0; c < b->nkids; ++c) free_subtree(fr, b->kids[c]);
  block_give(fr, b);
}

/* merge.c:85-107: the parent's volume is computed before the kids' gradients are
 * volume-weighted into it (a parent created and split in one pass has none yet).
 * Touches only b and reads its kids, so different parents may be done on different
 * threads. */
static void merge_average(const struct fvh_forest *fr, BRANCH *b) {
  const int neq = fr->neq;
  set_volume(b);
  for (int v = 0; v < neq; ++v) {
    b->el->S[v] = 0.0;
    b->el->SG[v][0] = b->el->SG[v][1] = b->el->SG[v][2] = 0.0;
    for (int c = 0; c < 8; ++c) {
      const BRANCH *k = b->kids[c];
      b->el->S[v] += k->el->S[v] / ((double)8);
      for (int d = 0; d < 3; ++d) b->el->SG[v][d] += k->el->SG[v][d] * k->cl->Vol / b->cl->Vol;
    }
  }
}

/* the kids of a merged parent leave the hash and the tree */
static void merge_unlink(struct fvh_forest *fr, BRANCH *b) {
  for (int c = 0; c < 8; ++c) {
    ht_remove(fr, b->kids[c]);
    free_subtree(fr, b->kids[c]);
    b->kids[c] = NULL;
  }
  b->nkids = 0;
  b->nsplit = 0;
}

/* appends the leaves of b's subtree in depth-first order; FVG_ERR_NOMEM leaves *out
 * valid (and owned by the caller) with what was collected */
static int collect_leaves(BRANCH *b, BRANCH ***out, int *n, int *cap) {
  if (b->nkids == 0) {
    if (*n == *cap) {
      int c = *cap ? 2 * *cap : 1024;
      BRANCH **q = (BRANCH **)realloc(*out, (size_t)c * sizeof(BRANCH *));
      if (!q) return FVG_ERR_NOMEM;
      *out = q;
      *cap = c;
    }
    (*out)[(*n)++] = b;
    return 0;
  }
  for (int c = 0; c < 8; ++c) {
    int rc = collect_leaves(b->kids[c], out, n, cap);
    if (rc) return rc;
  }
  return 0;
}

/* neighbour across face f of leaf b: fills neigtr/lfc_neigs/neigfc/nsfc and fc[f].bc */
static int resolve_face(struct fvh_forest *fr, BRANCH *b, int f) {
  int a = FV_FACE_AXIS[f], s = FV_FACE_SIGN[f];
  int n[3] = {IJK(b)[0], IJK(b)[1], IJK(b)[2]};
  int Na = ncell_axis(fr, b->level, a);
  n[a] += s;
  for (int k = 0; k < 4; ++k) b->neigtr[f][k] = NULL;
  b->neigag[f] = 0;
  if (n[a] < 0 || n[a] >= Na) {
    if (!fr->periodic[a]) {
      b->neigtr[f][0] = fr->glob;
      b->lfc_neigs[f] = (b->level == 1) ? 1 : 0;
      b->neigfc[f] = FV_FACE_OPP[f];
      b->nsfc[f] = 1;
      b->cl->fc[f].bc = fr->bc[FV_FACE_SIDE[f]];
      return 0;
    }
    n[a] = (n[a] + Na) % Na;
  }
  b->cl->fc[f].bc = 0;
  b->neigfc[f] = FV_FACE_OPP[f];
  BRANCH *nb = ht_find(fr, b->level, n);
  if (nb) {
    if (nb->nkids == 0) {
      b->neigtr[f][0] = nb;
      b->lfc_neigs[f] = 1;
    } else {
      b->lfc_neigs[f] = 4;
      for (int k = 0; k < 4; ++k) {
        int kid = FV_FACE_NODES[f][k] ^ (1 << a);
        BRANCH *kb = nb->kids[kid];
        if (kb->nkids) return FVG_ERR_STATE; /* 2:1 violated */
        b->neigtr[f][k] = kb;
      }
    }
  } else {
    if (b->level <= 1) return FVG_ERR_STATE;
    int m[3] = {n[0] >> 1, n[1] >> 1, n[2] >> 1};
    BRANCH *cb = ht_find(fr, b->level - 1, m);
    if (!cb || cb->nkids) return FVG_ERR_STATE; /* 2:1 violated */
    b->neigtr[f][0] = cb;
    b->lfc_neigs[f] = 1;
  }
  b->nsfc[f] = b->lfc_neigs[f] > 1 ? b->lfc_neigs[f] : 1;
  return 0;
}

/* growable list of branch pointers */
typedef struct {
  BRANCH **p;
  size_t n, cap;
} bvec;

static int bvec_push(bvec *v, BRANCH *b) {
  if (v->n == v->cap) {
    size_t cap = v->cap ? 2 * v->cap : 256;
    BRANCH **q = (BRANCH **)realloc(v->p, cap * sizeof(BRANCH *));
    if (!q) return FVG_ERR_NOMEM;
    v->p = q;
    v->cap = cap;
  }
  v->p[v->n++] = b;
  return 0;
}

/* leaf list in depth-first order with list links and indices.  The subtrees of a
 * frontier, widened in depth-first order until there is work for the threads, are
 * collected in parallel and concatenated in order. */
typedef struct {
  BRANCH **p;
  int n, cap, rc;
} lvec;

typedef struct {
  BRANCH *const *front;
  lvec *out;
  BRANCH **lv;
  BRANCH *glob;
  int n;
} relist_ctx;

static void relist_collect(void *v, int lo, int hi) {
  relist_ctx *r = (relist_ctx *)v;
  for (int t = lo; t < hi; ++t)
    r->out[t].rc = collect_leaves(r->front[t], &r->out[t].p, &r->out[t].n, &r->out[t].cap);
}

static void relist_link(void *v, int lo, int hi) {
  relist_ctx *r = (relist_ctx *)v;
  for (int i = lo; i < hi; ++i) {
    BRANCH *b = r->lv[i];
    IDX(b) = i;
    b->lprv = i ? r->lv[i - 1] : r->glob;
    b->lnxt = (i + 1 < r->n) ? r->lv[i + 1] : NULL;
  }
}

static int relist(struct fvh_forest *fr) {
  bvec a = {0}, b = {0};
  int rc = 0;
  for (int r = 0; r < fr->nroots && !rc; ++r) rc = bvec_push(&a, fr->roots[r]);
  while (!rc && a.n < 256) {
    int any = 0;
    b.n = 0;
    for (size_t q = 0; q < a.n && !rc; ++q) {
      BRANCH *x = a.p[q];
      if (x->nkids) {
        any = 1;
        for (int c = 0; c < 8 && !rc; ++c) rc = bvec_push(&b, x->kids[c]);
      } else {
        rc = bvec_push(&b, x);
      }
    }
    bvec t = a;
    a = b;
    b = t;
    if (!any) break;
  }
  free(b.p);
  const int nf = (int)a.n;
  lvec *out = rc ? NULL : (lvec *)calloc((size_t)(nf ? nf : 1), sizeof(lvec));
  if (!out) {
    free(a.p);
    return FVG_ERR_NOMEM;
  }
  relist_ctx ctx = {a.p, out, NULL, fr->glob, 0};
  par_for(nf, 4, relist_collect, &ctx);
  int64_t n = 0;
  for (int t = 0; t < nf; ++t) {
    n += out[t].n;
    if (out[t].rc) rc = out[t].rc;
  }
  if (!rc && n > INT32_MAX) rc = FVG_ERR_ARG;
  BRANCH **lv = rc ? NULL : (BRANCH **)realloc(fr->leaves, (size_t)(n ? n : 1) * sizeof(BRANCH *));
  if (lv) {
    fr->leaves = lv;
    int64_t o = 0;
    for (int t = 0; t < nf; ++t) {
      memcpy(lv + o, out[t].p, (size_t)out[t].n * sizeof(BRANCH *));
      o += out[t].n;
    }
  } else if (!rc) {
    rc = FVG_ERR_NOMEM;
  }
  for (int t = 0; t < nf; ++t) free(out[t].p);
  free(out);
  free(a.p);
  if (rc) return rc;
  fr->nleaves = (int)n;
  ctx.lv = fr->leaves;
  ctx.n = (int)n;
  par_for(n, 4096, relist_link, &ctx);
  fr->topo.locl = n ? fr->leaves[0] : NULL;
  fr->topo.nleaves = (int)n;
  fr->topo.tleaves = (int)n;
  return 0;
}

/* neighbours across all six faces, normals, areas, volume and centre of leaf b */
static int refresh_leaf(struct fvh_forest *fr, BRANCH *b) {
  for (int f = 0; f < 6; ++f) {
    int rc = resolve_face(fr, b, f);
    if (rc) return rc;
  }
  set_normals(b);
  set_props(b);
  return 0;
}

/* every leaf refreshed: after the forest is created */
/* refresh_leaf over a list of leaves, in parallel: each call reads the hash and its
 * neighbours' nodes, which do not change here, and writes only its own leaf */
typedef struct {
  struct fvh_forest *fr;
  BRANCH *const *lv;
  const int *idx; /* null: all of lv */
  int rc;
} refresh_ctx;

static void refresh_part(void *v, int lo, int hi) {
  refresh_ctx *r = (refresh_ctx *)v;
  for (int q = lo; q < hi; ++q) {
    int rc = refresh_leaf(r->fr, r->lv[r->idx ? r->idx[q] : q]);
    if (rc) __atomic_store_n(&r->rc, rc, __ATOMIC_RELAXED);
  }
}

static int rebuild(struct fvh_forest *fr) {
  int rc = relist(fr);
  if (rc) return rc;
  refresh_ctx ctx = {fr, fr->leaves, NULL, 0};
  par_for(fr->nleaves, 2048, refresh_part, &ctx);
  return ctx.rc;
}

/* position of the same-level cell across face f of node b, wrapped on periodic
 * axes; 0 when f lies on a non-periodic domain boundary */
static int across(const struct fvh_forest *fr, const BRANCH *b, int f, int n[3]) {
  int a = FV_FACE_AXIS[f];
  int Na = ncell_axis(fr, b->level, a);
  for (int d = 0; d < 3; ++d) n[d] = IJK(b)[d];
  n[a] += FV_FACE_SIGN[f];
  if (n[a] < 0 || n[a] >= Na) {
    if (!fr->periodic[a]) return 0;
    n[a] = (n[a] + Na) % Na;
  }
  return 1;
}

static void mark_subtree(const BRANCH *b, uint8_t *mark) {
  if (!b->nkids) { mark[IDX(b)] = 1; return; }
  for (int c = 0; c < 8; ++c) mark_subtree(b->kids[c], mark);
}

/* leaves of b's subtree that touch the cell across face f of b's same-level
 * neighbour: the kids on the side of b facing that cell, recursively */
static void mark_face(const BRANCH *b, int f, uint8_t *mark) {
  if (!b->nkids) { mark[IDX(b)] = 1; return; }
  int a = FV_FACE_AXIS[f];
  for (int k = 0; k < 4; ++k) mark_face(b->kids[FV_FACE_NODES[f][k] ^ (1 << a)], f, mark);
}

/* Leaf list, neighbours and geometry after an adaptation that split or merged the
 * nodes in chg.  Only leaves inside a changed node or face-adjacent to one can see
 * different neighbours (a leaf's neighbours are the same-level cell across each
 * face, its kids or the coarser cell), so only those are refreshed; the nodes of
 * every other leaf and of its neighbours are as before. */
static int rebuild_changed(struct fvh_forest *fr, const bvec *chg) {
  int rc = relist(fr);
  if (rc) return rc;
  uint8_t *mark = (uint8_t *)calloc((size_t)fr->nleaves + 1, 1);
  if (!mark) return FVG_ERR_NOMEM;
  for (size_t q = 0; q < chg->n; ++q) {
    const BRANCH *p = chg->p[q];
    mark_subtree(p, mark);
    for (int f = 0; f < 6; ++f) {
      int n[3];
      if (!across(fr, p, f, n)) continue;
      const BRANCH *nb = ht_find(fr, p->level, n);
      if (nb) {
        mark_face(nb, f, mark);
      } else if (p->level > 1) {
        int m[3] = {n[0] >> 1, n[1] >> 1, n[2] >> 1};
        const BRANCH *cb = ht_find(fr, p->level - 1, m);
        if (cb && !cb->nkids) mark[IDX(cb)] = 1;
      }
    }
  }
  int nm = 0;
  for (int i = 0; i < fr->nleaves; ++i) nm += mark[i];
  int *idx = (int *)malloc((size_t)(nm ? nm : 1) * sizeof(int));
  if (!idx) { free(mark); return FVG_ERR_NOMEM; }
  nm = 0;
  for (int i = 0; i < fr->nleaves; ++i)
    if (mark[i]) idx[nm++] = i;
  free(mark);
  refresh_ctx ctx = {fr, fr->leaves, idx, 0};
  par_for(nm, 2048, refresh_part, &ctx);
  free(idx);
  return ctx.rc;
}

/* ------------------------------------------------------------------ public */

int fvh_forest_create(const int nroot[3], const double x0[3], const double x1[3], int level,
                      const int bc[6], const int periodic[3], int neq, fvh_forest **out) {
  if (!out || !nroot || !x0 || !x1 || !bc || level < 0 || level > 16 || neq < 1 || neq > 8)
    return FVG_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (nroot[a] < 1 || !(x1[a] > x0[a])) return FVG_ERR_ARG;
  struct fvh_forest *fr = (struct fvh_forest *)calloc(1, sizeof(*fr));
  if (!fr) return FVG_ERR_NOMEM;
  fr->neq = neq;
  for (int a = 0; a < 3; ++a) {
    fr->nroot[a] = nroot[a];
    fr->periodic[a] = periodic ? (periodic[a] != 0) : 0;
    fr->x0[a] = x0[a];
    fr->x1[a] = x1[a];
  }
  for (int s = 0; s < 6; ++s) fr->bc[s] = bc[s];
  fr->run.con = &fr->con;
  fr->run.par = &fr->par;
  fr->run.topo = &fr->topo;
  fr->con.neq = neq;
  fr->con.neq_temp = neq;
  fr->con.nprimitiv = neq;
  fr->con.level = level;
  branch_layout(NULL, neq, &fr->bbytes);
  char *gb = block_take(fr);
  if (!gb) { fvh_forest_destroy(fr); return FVG_ERR_NOMEM; }
  fr->glob = branch_make(fr, gb);
  fr->glob->level = 0;
  fr->glob->root = -1;
  fr->glob->prnt = fr->glob;
  fr->glob->type = -1;
  fr->topo.glob = fr->glob;
  fr->nroots = nroot[0] * nroot[1] * nroot[2];
  fr->topo.ntrees = fr->nroots;
  fr->roots = (BRANCH **)calloc((size_t)fr->nroots, sizeof(BRANCH *));
  int r = 0;
  for (int k = 0; k < nroot[2]; ++k)
    for (int j = 0; j < nroot[1]; ++j)
      for (int i = 0; i < nroot[0]; ++i) {
        int ijk[3] = {i, j, k};
        char *m = block_take(fr);
        if (!m) { fvh_forest_destroy(fr); return FVG_ERR_NOMEM; }
        BRANCH *b = new_branch(fr, m, fr->glob, 1, ijk, r);
        fr->roots[r++] = b;
        int rc = ht_insert(fr, b);
        if (rc) { fvh_forest_destroy(fr); return rc; }
      }
  /* uniform refinement, breadth first */
  for (int l = 0; l < level; ++l) {
    int n = 0, cap = 0;
    BRANCH **lv = NULL;
    int rc = 0;
    for (int q = 0; q < fr->nroots && !rc; ++q) rc = collect_leaves(fr->roots[q], &lv, &n, &cap);
    if (!rc) rc = ht_reserve(fr, fr->hcount + (size_t)8 * n);
    for (int i = 0; i < n && !rc; ++i) rc = split_branch(fr, lv[i]);
    free(lv);
    if (rc) { fvh_forest_destroy(fr); return rc; }
  }
  int rc = rebuild(fr);
  if (rc) { fvh_forest_destroy(fr); return rc; }
  *out = fr;
  return FVG_OK;
}

void fvh_forest_destroy(fvh_forest *fr) {
  if (!fr) return;
  /* every branch lives in the mappings */
  for (int k = 0; k < fr->nslab; ++k) munmap(fr->slabs[k], FVH_SLAB);
  free(fr->slabs);
  free(fr->roots);
  free(fr->ht);
  free(fr->hk);
  free(fr->leaves);
  free(fr);
}

RUN *fvh_forest_run(fvh_forest *fr) { return fr ? &fr->run : NULL; }
int fvh_forest_nleaves(fvh_forest *fr) { return fr ? fr->nleaves : FVG_ERR_ARG; }

int fvh_forest_geometry(fvh_forest *fr, int nroot[3], double x0[3], double x1[3], int bc[6],
                        int periodic[3], int *neq) {
  if (!fr) return FVG_ERR_ARG;
  for (int a = 0; a < 3; ++a) {
    if (nroot) nroot[a] = fr->nroot[a];
    if (x0) x0[a] = fr->x0[a];
    if (x1) x1[a] = fr->x1[a];
    if (periodic) periodic[a] = fr->periodic[a];
  }
  for (int s = 0; s < 6 && bc; ++s) bc[s] = fr->bc[s];
  if (neq) *neq = fr->neq;
  return FVG_OK;
}

int fvh_forest_max_level(fvh_forest *fr) {
  if (!fr) return FVG_ERR_ARG;
  int m = 0;
  for (int i = 0; i < fr->nleaves; ++i)
    if (fr->leaves[i]->level > m) m = fr->leaves[i]->level;
  return m;
}

typedef struct {
  BRANCH *const *lv;
  int32_t *lev, *ijk;
} keys_ctx;

static void keys_part(void *v, int lo, int hi) {
  keys_ctx *k = (keys_ctx *)v;
  for (int i = lo; i < hi; ++i) {
    const BRANCH *b = k->lv[i];
    k->lev[i] = b->level;
    k->ijk[3 * i] = IJK(b)[0];
    k->ijk[3 * i + 1] = IJK(b)[1];
    k->ijk[3 * i + 2] = IJK(b)[2];
  }
}

BRANCH *const *fvh_forest_leaves(fvh_forest *fr) { return fr ? (BRANCH *const *)fr->leaves : NULL; }

int fvh_forest_remesh_device(fvh_forest *fr, fvg_ctx *ctx, int carry_state) {
  if (!fr || !ctx) return FVG_ERR_ARG;
  const int n = fr->nleaves;
  if (n <= 0) return FVG_ERR_ARG;
  int32_t *lev = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *ijk = (int32_t *)malloc((size_t)3 * n * sizeof(int32_t));
  double *tab = (double *)malloc(sizeof(double) * 32 * (3 + 18 + 3 + 1));
  if (!lev || !ijk || !tab) { free(lev); free(ijk); free(tab); return FVG_ERR_NOMEM; }
  keys_ctx kc = {fr->leaves, lev, ijk};
  par_for(n, 8192, keys_part, &kc);
  int32_t own[FVG_REMESH_OWNERS];
  int rc = fvg_remesh_begin(ctx, n, lev, ijk, fr->nroot, fr->periodic, fr->bc, own);
  free(lev);
  free(ijk);
  double *h = tab, *dist = tab + 32 * 3, *area = dist + 32 * 18, *vol = area + 32 * 3;
  int32_t nl = 0;
  if (!rc) rc = fvh_tables_from_owners(fr->leaves, n, own, h, dist, area, vol, &nl);
  if (!rc) rc = fvg_remesh_end(ctx, nl, h, dist, area, vol, carry_state);
  free(tab);
  return rc;
}

int fvh_forest_leaf_info(fvh_forest *fr, double *xc, int32_t *level) {
  if (!fr) return FVG_ERR_ARG;
  for (int i = 0; i < fr->nleaves; ++i) {
    const BRANCH *b = fr->leaves[i];
    if (xc) {
      xc[3 * i] = b->cl->xc;
      xc[3 * i + 1] = b->cl->yc;
      xc[3 * i + 2] = b->cl->zc;
    }
    if (level) level[i] = b->level;
  }
  return FVG_OK;
}

/* var-major layouts as FVG_FIELD_* */
static int field_io(fvh_forest *fr, int which, double *buf, int to_leaves) {
  if (!fr || !buf) return FVG_ERR_ARG;
  const int N = fr->nleaves, neq = fr->neq;
  for (int i = 0; i < N; ++i) {
    LEAF *el = fr->leaves[i]->el;
    switch (which) {
      case FVG_FIELD_S:
      case FVG_FIELD_SN:
      case FVG_FIELD_RHS: {
        double *p = which == FVG_FIELD_S ? el->S : (which == FVG_FIELD_SN ? el->SN : el->RHS);
        for (int v = 0; v < neq; ++v) {
          if (to_leaves) p[v] = buf[(size_t)v * N + i];
          else buf[(size_t)v * N + i] = p[v];
        }
        break;
      }
      case FVG_FIELD_SF:
        for (int f = 0; f < 6; ++f)
          for (int v = 0; v < neq; ++v) {
            size_t o = ((size_t)f * neq + v) * N + i;
            if (to_leaves) el->SF[f][v] = buf[o];
            else buf[o] = el->SF[f][v];
          }
        break;
      case FVG_FIELD_SG:
        for (int v = 0; v <= neq; ++v)
          for (int d = 0; d < 3; ++d) {
            size_t o = ((size_t)v * 3 + d) * N + i;
            if (to_leaves) el->SG[v][d] = buf[o];
            else buf[o] = el->SG[v][d];
          }
        break;
      default:
        return FVG_ERR_ARG;
    }
  }
  return FVG_OK;
}

int fvh_forest_set_field(fvh_forest *fr, int which, const double *src) {
  return field_io(fr, which, (double *)src, 1);
}
int fvh_forest_get_field(fvh_forest *fr, int which, double *dst) {
  return field_io(fr, which, dst, 0);
}

/* a merge of parent p keeps 2:1 only if no kid has a finer face neighbour */
static int merge_allowed(const struct fvh_forest *fr, const BRANCH *p) {
  for (int c = 0; c < 8; ++c) {
    const BRANCH *k = p->kids[c];
    for (int f = 0; f < 6; ++f) {
      int n[3];
      if (!across(fr, k, f, n)) continue;
      const BRANCH *nb = ht_find(fr, k->level, n);
      if (nb && nb->nkids) return 0;
    }
  }
  return 1;
}

/* Split leaf b and queue the leaves that the split leaves two levels coarser than
 * a face neighbour: the leaves one level coarser than b across its faces.  In a
 * 2:1-balanced forest those are the only new violations, and each of them must be
 * split in any balanced refinement, so working the queue empty gives the minimal
 * balanced forest whatever the order. */
static int split_queue(struct fvh_forest *fr, BRANCH *b, bvec *chg, bvec *work) {
  int rc = split_branch(fr, b);
  if (!rc) rc = bvec_push(chg, b);
  for (int f = 0; f < 6 && !rc && b->level > 1; ++f) {
    int n[3];
    if (!across(fr, b, f, n) || ht_find(fr, b->level, n)) continue;
    int m[3] = {n[0] >> 1, n[1] >> 1, n[2] >> 1};
    BRANCH *c = ht_find(fr, b->level - 1, m);
    if (c && !c->nkids) rc = bvec_push(work, c);
  }
  return rc;
}

typedef struct {
  const struct fvh_forest *fr;
  BRANCH *const *leaves;
  char *const *blk; /* 8 per leaf */
} kids_ctx;

static void kids_part(void *v, int lo, int hi) {
  kids_ctx *k = (kids_ctx *)v;
  for (int q = lo; q < hi; ++q) split_kids(k->fr, k->leaves[q], k->blk + 8 * (size_t)q);
}

/* the per-leaf and per-candidate passes of an adaptation, each on the thread pool: every
 * item writes only its own entries and nodes */
typedef struct {
  const struct fvh_forest *fr;
  BRANCH *const *lv;
  const int8_t *flag;
  int max_level;
  uint8_t *split, *first; /* per leaf */
  BRANCH *const *cand;
  uint8_t *keep; /* per candidate */
} pass_ctx;

static void flag_part(void *v, int lo, int hi) {
  pass_ctx *p = (pass_ctx *)v;
  for (int i = lo; i < hi; ++i) {
    BRANCH *b = p->lv[i];
    b->split = (p->flag[i] > 0 && b->level < p->max_level);
    b->merge = (p->flag[i] < 0);
    p->split[i] = (uint8_t)b->split;
  }
}

static void cand_part(void *v, int lo, int hi) {
  pass_ctx *p = (pass_ctx *)v;
  for (int i = lo; i < hi; ++i) {
    const BRANCH *b = p->lv[i];
    int ok = b->level > 1 && b->prnt->kids[0] == b;
    for (int c = 0; c < 8 && ok; ++c) ok = (b->prnt->kids[c]->nkids == 0 && b->prnt->kids[c]->merge);
    p->first[i] = (uint8_t)ok;
  }
}

static void merge_decide_part(void *v, int lo, int hi) {
  pass_ctx *p = (pass_ctx *)v;
  for (int q = lo; q < hi; ++q) {
    const BRANCH *b = p->cand[q];
    int ok = 1;
    for (int c = 0; c < 8 && ok; ++c) ok = (b->kids[c]->nkids == 0);
    p->keep[q] = (uint8_t)(ok && merge_allowed(p->fr, b));
  }
}

static void merge_average_part(void *v, int lo, int hi) {
  pass_ctx *p = (pass_ctx *)v;
  for (int q = lo; q < hi; ++q) merge_average(p->fr, p->cand[q]);
}

static void clear_part(void *v, int lo, int hi) {
  pass_ctx *p = (pass_ctx *)v;
  for (int i = lo; i < hi; ++i) p->lv[i]->split = p->lv[i]->merge = 0;
}

int fvh_forest_adapt(fvh_forest *fr, const int8_t *flag, int max_level) {
  if (!fr || !flag) return FVG_ERR_ARG;
  const int N = fr->nleaves;
  BRANCH **lv = fr->leaves;
  /* per leaf: flagged split, and first kid of a merge candidate (a parent on the old
   * tree whose 8 kids are leaves flagged -1) */
  uint8_t *sm = (uint8_t *)malloc((size_t)2 * N + 1);
  BRANCH **cand = (BRANCH **)malloc((size_t)(N / 8 + 1) * sizeof(BRANCH *));
  if (!sm || !cand) { free(sm); free(cand); return FVG_ERR_NOMEM; }
  pass_ctx pc = {fr, lv, flag, max_level, sm, sm + N, NULL, NULL};
  par_for(N, 8192, flag_part, &pc);
  par_for(N, 8192, cand_part, &pc);
  int ncand = 0;
  size_t nflag = 0;
  for (int i = 0; i < N; ++i) {
    nflag += sm[i];
    if (sm[N + i]) cand[ncand++] = lv[i]->prnt;
  }
  free(sm);
  /* room for the flagged splits; balance splits grow the hash as they go */
  bvec chg = {0}, work = {0};
  int rc = ht_reserve(fr, fr->hcount + 8 * nflag);
  /* the flagged leaves' kids are made on several threads, then linked in list order */
  BRANCH **sp = rc ? NULL : (BRANCH **)malloc((nflag ? nflag : 1) * sizeof(BRANCH *));
  if (!rc && !sp) rc = FVG_ERR_NOMEM;
  if (!rc) {
    size_t q = 0;
    for (int i = 0; i < N; ++i)
      if (lv[i]->split) sp[q++] = lv[i];
    /* the blocks are taken here, in order; the threads only fill them */
    char **blk = (char **)malloc((8 * nflag + 1) * sizeof(char *));
    size_t nb = 0;
    if (!blk) rc = FVG_ERR_NOMEM;
    for (; nb < 8 * nflag && !rc; ++nb)
      if (!(blk[nb] = block_take(fr))) rc = FVG_ERR_NOMEM;
    if (rc) {
      for (size_t k = 0; blk && k + 1 < nb; ++k) block_give(fr, blk[k]);
      nflag = 0; /* nothing made: nothing to link or drop */
    } else {
      kids_ctx kc = {fr, sp, blk};
      par_for((int)nflag, 256, kids_part, &kc);
    }
    free(blk);
    for (size_t k = 0; k < nflag && !rc; ++k) rc = split_queue(fr, sp[k], &chg, &work);
    /* after a failure, kids made but not linked are dropped */
    for (size_t k = 0; k < nflag && rc; ++k)
      if (!sp[k]->nkids)
        for (int c = 0; c < 8; ++c) {
          block_give(fr, sp[k]->kids[c]);
          sp[k]->kids[c] = NULL;
        }
  }
  free(sp);
  /* 2:1 face balance by further splits */
  while (!rc && work.n) {
    BRANCH *c = work.p[--work.n];
    if (!c->nkids) rc = split_queue(fr, c, &chg, &work);
  }
  free(work.p);
  /* after a failure some splits may be linked already: relist the leaves so that the
   * forest stays consistent (balanced or not) for the caller's next step */
  if (rc) { free(cand); free(chg.p); (void)rebuild(fr); return rc; }
  /* decide every merge on the same (post-split) tree, then apply: the decisions and the
   * means on the thread pool, the unlinking in order */
  int nm = 0;
  uint8_t *keep = rc ? NULL : (uint8_t *)malloc((size_t)ncand + 1);
  if (!rc && !keep) rc = FVG_ERR_NOMEM;
  if (!rc) {
    pc.cand = cand;
    pc.keep = keep;
    par_for(ncand, 1024, merge_decide_part, &pc);
    for (int q = 0; q < ncand; ++q)
      if (keep[q]) cand[nm++] = cand[q];
    par_for(nm, 1024, merge_average_part, &pc);
    for (int q = 0; q < nm && !rc; ++q) {
      merge_unlink(fr, cand[q]);
      rc = bvec_push(&chg, cand[q]);
    }
  }
  free(keep);
  free(cand);
  /* fr->leaves still lists freed kids of merged parents: rebuild before any use */
  if (!rc) rc = rebuild_changed(fr, &chg);
  free(chg.p);
  if (rc) { (void)rebuild(fr); return rc; }
  pc.lv = fr->leaves;
  par_for(fr->nleaves, 8192, clear_part, &pc);
  return fr->nleaves;
}

/* ------------------------------------------------------------------ restart */

/* Binary restart of this library: the forest's description, the leaf levels in list
 * order (a depth-first level sequence fixes the tree over given roots) and S var-major,
 * all native-endian.  The reference has no binary restart; its text layout follows. */
#define FVH_MAGIC "FVH-RST1"

typedef struct {
  char magic[8];
  int32_t neq, nroot[3], bc[6], periodic[3], tstep;
  double x0[3], x1[3], time;
  int64_t nleaves;
} fvh_rst_head;

int fvh_forest_save(fvh_forest *fr, const char *path, int32_t tstep, double time) {
  if (!fr || !path) return FVG_ERR_ARG;
  fvh_rst_head h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, FVH_MAGIC, 8);
  h.neq = fr->neq;
  for (int a = 0; a < 3; ++a) {
    h.nroot[a] = fr->nroot[a];
    h.periodic[a] = fr->periodic[a];
    h.x0[a] = fr->x0[a];
    h.x1[a] = fr->x1[a];
  }
  for (int s = 0; s < 6; ++s) h.bc[s] = fr->bc[s];
  h.tstep = tstep;
  h.time = time;
  h.nleaves = fr->nleaves;
  const size_t N = (size_t)fr->nleaves;
  uint8_t *lev = (uint8_t *)malloc(N ? N : 1);
  double *S = (double *)malloc((N ? N : 1) * (size_t)fr->neq * sizeof(double));
  if (!lev || !S) { free(lev); free(S); return FVG_ERR_NOMEM; }
  for (size_t i = 0; i < N; ++i) lev[i] = (uint8_t)fr->leaves[i]->level;
  field_io(fr, FVG_FIELD_S, S, 0);
  FILE *fp = fopen(path, "wb");
  int rc = fp ? 0 : FVG_ERR_IO;
  if (!rc && (fwrite(&h, sizeof(h), 1, fp) != 1 || fwrite(lev, 1, N, fp) != N ||
              fwrite(S, sizeof(double), N * fr->neq, fp) != N * fr->neq))
    rc = FVG_ERR_IO;
  if (fp && fclose(fp) != 0) rc = FVG_ERR_IO;
  free(lev);
  free(S);
  return rc;
}

/* rebuild the subtree of node b from the depth-first leaf levels lev[*pos..n) */
static int grow_from_levels(struct fvh_forest *fr, BRANCH *b, const uint8_t *lev, size_t n,
                            size_t *pos) {
  if (*pos >= n) return FVG_ERR_ARG;
  if (lev[*pos] == b->level) { ++*pos; return 0; }
  if (lev[*pos] < b->level || b->level >= 30) return FVG_ERR_ARG;
  int rc = split_branch(fr, b);
  for (int c = 0; c < 8 && !rc; ++c) rc = grow_from_levels(fr, b->kids[c], lev, n, pos);
  return rc;
}

int fvh_forest_load(const char *path, fvh_forest **out, int32_t *tstep, double *time) {
  if (!path || !out) return FVG_ERR_ARG;
  FILE *fp = fopen(path, "rb");
  if (!fp) return FVG_ERR_IO;
  fvh_rst_head h;
  uint8_t *lev = NULL;
  double *S = NULL;
  fvh_forest *fr = NULL;
  int rc = fread(&h, sizeof(h), 1, fp) == 1 ? 0 : FVG_ERR_IO;
  if (!rc && (memcmp(h.magic, FVH_MAGIC, 8) != 0 || h.nleaves < 1 || h.nleaves > INT32_MAX ||
              h.neq < 1 || h.neq > 8))
    rc = FVG_ERR_ARG;
  const size_t N = rc ? 0 : (size_t)h.nleaves;
  if (!rc) {
    lev = (uint8_t *)malloc(N);
    S = (double *)malloc(N * (size_t)h.neq * sizeof(double));
    if (!lev || !S) rc = FVG_ERR_NOMEM;
  }
  if (!rc && (fread(lev, 1, N, fp) != N || fread(S, sizeof(double), N * h.neq, fp) != N * h.neq))
    rc = FVG_ERR_IO;
  fclose(fp);
  if (!rc) rc = fvh_forest_create(h.nroot, h.x0, h.x1, 0, h.bc, h.periodic, h.neq, &fr);
  size_t pos = 0;
  for (int r = 0; !rc && r < fr->nroots; ++r) rc = grow_from_levels(fr, fr->roots[r], lev, N, &pos);
  if (!rc && pos != N) rc = FVG_ERR_ARG;
  if (!rc) rc = rebuild(fr);
  if (!rc) rc = field_io(fr, FVG_FIELD_S, S, 1);
  free(lev);
  free(S);
  if (rc) { fvh_forest_destroy(fr); return rc; }
  if (tstep) *tstep = h.tstep;
  if (time) *time = h.time;
  *out = fr;
  return FVG_OK;
}

/* The reference's text restart (restart_save.c:17-45): a header line
 * "tstep time nleaves ntrees neq" and per leaf in list order "root tag  S0 ... " with
 * each value as " %15.10e ".  The tag is the leaf's kid path (tagaddress.c:19-27), so
 * levels above 10 cannot be written. */
int fvh_forest_save_text(fvh_forest *fr, const char *path, int32_t tstep, double time) {
  if (!fr || !path) return FVG_ERR_ARG;
  for (int i = 0; i < fr->nleaves; ++i)
    if (fr->leaves[i]->tag < 0) return FVG_ERR_UNSUPPORTED;
  FILE *fp = fopen(path, "w");
  if (!fp) return FVG_ERR_IO;
  fprintf(fp, "%d %15.10e %d %d %d \n", tstep, time, fr->nleaves, fr->nroots, fr->neq);
  for (int i = 0; i < fr->nleaves; ++i) {
    const BRANCH *b = fr->leaves[i];
    fprintf(fp, "%d %d ", b->root, b->tag);
    for (int v = 0; v < fr->neq; ++v) fprintf(fp, " %15.10e ", b->el->S[v]);
    fprintf(fp, "\n");
  }
  return (ferror(fp) | fclose(fp)) ? FVG_ERR_IO : FVG_OK;
}

/* restart_load.c:30-58 with spawntree.c:40-54,112-125: each leaf line names a root and
 * a tag whose decimal digits, least significant first, are 1 + the kid index at levels
 * 2, 3, ...; the path is split open as needed and the leaf takes the listed state.  The
 * forest must be bare roots of the geometry the restart was written on, and every leaf
 * of the result must be listed exactly once. */
int fvh_forest_load_text(fvh_forest *fr, const char *path, int32_t *tstep, double *time) {
  if (!fr || !path) return FVG_ERR_ARG;
  for (int r = 0; r < fr->nroots; ++r)
    if (fr->roots[r]->nkids) return FVG_ERR_STATE;
  FILE *fp = fopen(path, "r");
  if (!fp) return FVG_ERR_IO;
  int ts = 0, nl = 0, nt = 0, nv = 0;
  double tm = 0.0;
  int rc = fscanf(fp, "%d %lf %d %d %d", &ts, &tm, &nl, &nt, &nv) == 5 ? 0 : FVG_ERR_IO;
  if (!rc && (nt != fr->nroots || nv != fr->neq || nl < 1)) rc = FVG_ERR_ARG;
  bvec got = {0};
  for (int l = 0; l < nl && !rc; ++l) {
    int root;
    long long tag;
    if (fscanf(fp, "%d %lld", &root, &tag) != 2) { rc = FVG_ERR_IO; break; }
    if (root < 0 || root >= fr->nroots || tag < 0) { rc = FVG_ERR_ARG; break; }
    BRANCH *b = fr->roots[root];
    for (long long t = tag; t > 0 && !rc; t /= 10) {
      int d = (int)(t % 10);
      if (d < 1 || d > 8 || b->level >= 30) { rc = FVG_ERR_ARG; break; }
      rc = split_branch(fr, b);
      if (!rc) b = b->kids[d - 1];
    }
    if (!rc && b->nkids) rc = FVG_ERR_ARG;  /* a listed cell that is not a leaf */
    for (int v = 0; v < nv && !rc; ++v)
      if (fscanf(fp, "%lf", &b->el->S[v]) != 1) rc = FVG_ERR_IO;
    if (!rc) rc = bvec_push(&got, b);
  }
  fclose(fp);
  if (!rc) rc = rebuild(fr);
  if (!rc && fr->nleaves != nl) rc = FVG_ERR_ARG;
  /* every leaf listed once: list positions of the listed leaves are a permutation */
  if (!rc) {
    uint8_t *seen = (uint8_t *)calloc((size_t)nl, 1);
    if (!seen) rc = FVG_ERR_NOMEM;
    for (size_t q = 0; q < got.n && !rc; ++q) {
      int i = IDX(got.p[q]);
      if (got.p[q]->nkids || i < 0 || i >= nl || seen[i]) rc = FVG_ERR_ARG;
      else seen[i] = 1;
    }
    free(seen);
  }
  free(got.p);
  if (rc) return rc;
  if (tstep) *tstep = ts;
  if (time) *time = tm;
  return FVG_OK;
}
