/*
 * fv_strdata.h — host data model of the octree finite-volume solver.
 *
 * This header declares the subset of the reference's strdata.h types that the
 * RHS/RK hot path reads and writes, with the same type names, field names and
 * meanings (strdata.h:62-540), so that host code written against the reference
 * (RHS(run), RK(irk, run), faceinterp(run), gradcalc(run)) compiles against it
 * unchanged.  MPI, METIS and PETSc members are omitted: the GPU path has no
 * halo exchange and no partitioner.
 *
 * Fields marked "extension" do not exist in the reference; they carry the
 * octree's integer addresses and the periodic box used by this build.
 *
 * Ownership: every struct reachable from a RUN built by fvh_forest_create()
 * is owned by that forest handle and freed by fvh_forest_destroy().
 */
#ifndef FV_STRDATA_H
#define FV_STRDATA_H

#ifdef __cplusplus
extern "C" {
#endif

/* strdata.h:62-67 */
typedef struct ND {
  int num;
  double x;
  double y;
  double z;
} ND;

/* strdata.h:69-74; bc: 0 interior (or periodic), 1 inlet, 2 outflow/transmissive,
 * 3 reflective, 4 no-slip wall (3 and 4 are undefined in the reference: bc.c:38,42) */
typedef struct FC {
  int num;
  int type;
  int nnds;
  int bc;
} FC;

/* strdata.h:76-93 — per-leaf state.  S/SN/RHS hold neq doubles, SF[6][neq] the
 * MUSCL face states, SG[neq+1][3] the Green-Gauss gradients (conserved + p). */
typedef struct LEAF {
  int gnum;
  double *RHS;
  double *S;
  double *SN;
  double **SF;
  double **SG;
  double crith;
  double **flux_face;
  int *flux_flag;
  int inner;
} LEAF;

/* strdata.h:150-168 — per-leaf geometry. */
typedef struct CELL {
  double *nx;
  double *ny;
  double *nz;
  double *nxt1;
  double *nyt1;
  double *nzt1;
  double *nxt2;
  double *nyt2;
  double *nzt2;
  double Vol;
  double **Area;
  double xc;
  double yc;
  double zc;
  struct ND *nd;
  struct FC *fc;
} CELL;

/* strdata.h:104-148 — tree node (branch or leaf). */
typedef struct BRANCH {
  int igroup;
  int part;
  int oldpart;
  int level;          /* roots are level 1, glob (outside the domain) level 0 */
  int type;           /* 0 = hexahedron */
  int nkeens;
  int nkids;
  int nlnd;
  int nlfc;
  int tot_neigs;
  int *adrs;
  int split;
  int merge;
  int tag;
  int hangn;
  struct BRANCH *prnt;
  int root;
  struct BRANCH *next;
  struct BRANCH *lnxt;   /* leaf list, depth-first = Morton order within a root */
  struct BRANCH *lprv;
  struct BRANCH ***neigtr;  /* neigtr[f][k]: k-th neighbour across face f */
  int **neignd;
  int *neigfc;              /* neighbour's face id facing this leaf */
  int *neigag;
  int *lfc_neigs;           /* 1 same-or-coarser (or root boundary), 4 finer, 0 refined-cell boundary */
  int *nsfc;                /* max(1, lfc_neigs) */
  int **ipartbound;
  struct BRANCH **keen;
  int *keenfc;
  int *keenfcangle;
  int fcqsplit;
  int nsplit;
  int ctfc;
  struct BRANCH **kids;
  struct LEAF *el;
  struct CELL *cl;
  /* extension: integer address at this level (root grid index scaled by 2^(level-1)) */
  int fv_ijk[3];
  /* extension: position of the leaf in the leaf list (set by fvh_forest_rebuild) */
  int fv_index;
} BRANCH;

/* strdata.h:170-197 (subset) */
typedef struct FOREST {
  int partitioned;
  int ntrees;
  int ndepth;
  int nleaves;
  int tleaves;
  struct BRANCH *locl;   /* head of the leaf list */
  struct BRANCH *glob;   /* the branch outside the domain (level 0, root -1) */
  /* extension: periodic lengths per axis (0 = not periodic) */
  double fv_period[3];
  /* extension: number of roots per axis and the box */
  int fv_nroot[3];
  double fv_x0[3];
  double fv_x1[3];
} FOREST;

/* strdata.h:199-210 (subset) */
typedef struct PAR {
  double matervisc[10];
  int bcoverwrt[10];
  double materpinf[10];
  double matermush[10];
  double matergama[10];
  double materrini[10];
  double materpini[10];
} PAR;

/* strdata.h:212-358 (subset touched by the hot path) */
typedef struct CON {
  int level;
  int ns;
  int turb;
  int i_case;
  int nprimitiv;
  double ymin;
  int model;
  int multimat;
  int nstep;
  double dt;
  int eos;
  int verbose;
  int halfdt;
  int limiter;
  int istep;
  int tstep;
  double time;
  int rank;
  int size;
  int neq;
  int neq_temp;
  int ischeme;
  int iprimtv;
  int method;
} CON;

/* strdata.h:405-540 (subset) */
typedef struct RUN {
  struct CON *con;
  struct PAR *par;
  struct FOREST *topo;
} RUN;

/* Reference hot-path entry points (strdata.h:642,654,656,675). */
void RHS(struct RUN *run);
void RK(int irk, RUN *run);
void faceinterp(RUN *run);
void gradcalc(struct RUN *run);

#ifdef __cplusplus
}
#endif

#endif /* FV_STRDATA_H */
