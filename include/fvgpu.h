/*
 * fvgpu.h — C ABI of the MI355X (gfx950) right-hand-side + Runge-Kutta path.
 *
 * Three groups of entry points live in libfv140823.so:
 *
 *  1. fvh_*  host octree + flattener.  The forest is a tree of the reference's
 *     BRANCH/LEAF/CELL structs (fv_strdata.h); fvh_flatten_run() turns the
 *     leaf list of any RUN into the struct-of-arrays mesh the kernels read.
 *  2. fvg_*  device context: state upload/download and the kernels that
 *     replace faceinterp() (faceinterp.c:21), gradcalc() (gradcalc.c:9),
 *     RHS() (RHS.c:3 -> residual_faces_AMR.c:10) and RK() (RK.c:3).
 *  3. RHS/RK/faceinterp/gradcalc with the reference signatures (declared in
 *     fv_strdata.h), bound to a device context by fv_gpu_attach().
 *
 * All functions return an int status (FVG_OK or a negative FVG_ERR_*), except
 * where stated.  A context owns one HIP stream created at default priority;
 * calls on one context must not overlap from several host threads.
 */
#ifndef FVGPU_H
#define FVGPU_H

#include <stdint.h>
#include "fv_strdata.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FVG_API_VERSION 1

enum {
  FVG_OK = 0,
  FVG_ERR_ARG = -1,         /* bad argument or size */
  FVG_ERR_HIP = -2,         /* a HIP runtime call failed */
  FVG_ERR_STATE = -3,       /* call out of order (e.g. no mesh yet) */
  FVG_ERR_NONFINITE = -4,   /* NaN/Inf in RHS or state; first bad cell via fvg_first_bad_cell */
  FVG_ERR_UNSUPPORTED = -5, /* option with no defined behaviour (bc 3/4, model 4, ...) */
  FVG_ERR_NOMEM = -6,
  FVG_ERR_IO = -7           /* a file could not be opened, read or written */
};

/* precision forms */
enum { FVG_PLAIN = 0, FVG_MIXED = 1 };
/* limiters: minmod is the reference's hard-wired one (reconctruct_V3.c:19-30) */
enum { FVG_MINMOD = 0, FVG_VANALBADA = 1 };
/* wave speeds of the HLLC solver: Davis, the reference's (compute_wave_speeds.c:53-54), or
 * the primitive-variable (PVRS) pressure estimate (Toro, Riemann Solvers, 3rd ed., 10.5.2) */
enum { FVG_WAVES_DAVIS = 0, FVG_WAVES_PVRS = 1 };
/* fields for fvg_get_field / fvg_device_ptr */
enum {
  FVG_FIELD_S = 0,    /* [neq][ncell] f64 */
  FVG_FIELD_SN = 1,   /* [neq][ncell] f64 */
  FVG_FIELD_RHS = 2,  /* [neq][ncell] f64, already multiplied by -Area, not divided by Vol */
  FVG_FIELD_SF = 3,   /* [6][neq][ncell] (f64 plain / f32 mixed on device; f64 on copy-out) */
  FVG_FIELD_SG = 4,   /* [neq+1][3][ncell] f64 (gradcalc output) */
  FVG_FIELD_FACEFLUX = 5, /* [neq][nface] face flux of the owner side (f64 / f32) */
  /* [neq][ncell] f64, read only: after fvg_step / fvg_steps, the state the last stage of
   * the last step read, i.e. the input whose residual FVG_FIELD_RHS holds.  FVG_ERR_STATE
   * once anything else has written the state since. */
  FVG_FIELD_STAGE_IN = 6
};

typedef struct fvg_params {
  int model;      /* 1 ideal/stiffened gas (neq 5); 2 barotropic liquid-vapour (neq 4);
                     3 barotropic two-phase mixture with mass fraction (neq 5) */
  int ns;         /* 1: viscous fluxes (residual_faces_AMR.c:129) */
  int ischeme;    /* 1: first order face states (facevalues.c); 2: MUSCL (faceinterp.c) */
  int limiter;    /* FVG_MINMOD | FVG_VANALBADA */
  int precision;  /* FVG_PLAIN | FVG_MIXED */
  int visc_work;  /* the reference's con->eos == 1: subtract u.tau from row 4 (flux_viscous.c:227-232);
                     default 0, the value the never-set flag reads in the reference */
  double gamma;   /* par->matergama[0] */
  double pinf;    /* par->materpinf[0] */
  double mu;      /* par->matervisc[0] */
  double ymin;    /* model 3 mass-fraction clip at ymin/1000 (facercnstr.c:43-44) */
  /* model 2 constants (barotropic_eos.c:8-12) */
  double baro_nexp, baro_beta, baro_cbar, baro_psat, baro_rsat;
  /* model 3 constants (barotropic_eos.c:40-46) */
  double mix_RG, mix_Tref, mix_rsat, mix_psat, mix_cliq, mix_cmix;
  double inlet[5];        /* conserved ghost state of bc type 1 (bc_inlet.c) */
  double vanalbada_eps;   /* epsilon of the smooth van Albada limiter */
  /* Options beyond the reference's behaviour; 0 keeps it. */
  int iprimtv;      /* 1: MUSCL on the primitive variables (rho, u, v, w, p | Y) of the reference's
                       iprimtv = 1 path (deltavalues.c:16-19, reconctruct_V3.c:37-51, prim2conc.c),
                       with the state's own density and the kinetic energy from its velocities
                       (cons2primtv.c:10-18 reads a stale density and squares the momenta) */
  int wave_speeds;  /* FVG_WAVES_DAVIS | FVG_WAVES_PVRS (model 1) */
  double prandtl;   /* > 0: heat conduction in the energy row of the viscous flux (model 1, ns),
                       q = -mu gamma / ((gamma - 1) Pr) grad(p / rho), the gradient a Green-Gauss one
                       of p / rho averaged over the face's two cells; 0: none (flux_viscous.c has
                       no heat flux) */
} fvg_params;

/* Flattened mesh (all arrays host memory, read by fvg_set_mesh). */
typedef struct fvg_mesh {
  int32_t ncell;
  int32_t nface;          /* unique faces */
  int32_t nface_axis[3];  /* faces 0..nface_axis[0]-1 are x faces, then y, then z (incl. boundary) */
  int32_t nquad;          /* coarse faces with 4 finer neighbours */
  int32_t nlevel;         /* rows in the per-level tables */
  int32_t neq;
  const uint32_t *cflag;  /* [ncell] bits 2f..2f+1: class of face f (0 same level, 1 finer (quad),
                             2 coarser, 3 domain boundary); bits 16-23: level */
  const int32_t *nbr;     /* [6][ncell]: same/coarser -> neighbour cell; finer -> quad index;
                             boundary -> -1-bc */
  const int32_t *quad_nbr;  /* [nquad][4] fine neighbours in sub-face order */
  const int32_t *face_L;  /* [nface] owner cell */
  const int32_t *face_R;  /* [nface] other cell (= owner for boundary faces) */
  const uint32_t *face_meta;/* [nface] bits 0-2 owner face id, 3-5 other face id,
                               6-7 bc (0 interior, 1 inlet, 2 outflow), 8-15 level of the finer side */
  const int32_t *gslot;   /* [6][ncell] (face<<1)|other_side; class finer: first of 4 consecutive
                             owned faces; -1: no flux through this face */
  const double *h;        /* [nlevel][3] cell size per level and axis */
  const double *dist;     /* [nlevel][3][3][2] (d_cf, d_cc) per level, axis, class (distance.c) */
  const double *area;     /* [nlevel][3] face area per level of the finer side and axis */
  const double *vol;      /* [nlevel] cell volume */
  /* Within each axis the first nface_reg[a] faces are regular: interior, same level, and
   * both cells have same-level neighbours on their far sides along the axis (a MUSCL
   * stencil of four cells of one level).  For them face_LL / face_RR hold the far-side
   * neighbours of face_L / face_R; for the others, -1. */
  int32_t nface_reg[3];
  const int32_t *face_LL; /* [nface] */
  const int32_t *face_RR; /* [nface] */
} fvg_mesh;

typedef struct fvg_ctx fvg_ctx;
typedef struct fvh_forest fvh_forest;
typedef struct fvh_flat fvh_flat;

/* ---------------------------------------------------------------- params */
/* Fill defaults: gamma 1.4, pinf 0, mu 0, the reference's EOS constants, minmod, plain. */
void fvg_params_default(fvg_params *p, int model);
int fvg_neq(int model);     /* 5, 4, 5; -1 unsupported */

/* ---------------------------------------------------------------- host forest */
/* Build a forest of nroot[0]*nroot[1]*nroot[2] hexahedral roots over the box [x0,x1],
 * refined uniformly by `level` (level 0 = the roots).  bc[6] are the side types in the
 * reference's order -x,+x,-y,+y,-z,+z (inputset.c:99-104); periodic[a]!=0 pairs the two
 * sides of axis a.  neq: conserved variables per leaf. */
int fvh_forest_create(const int nroot[3], const double x0[3], const double x1[3], int level,
                      const int bc[6], const int periodic[3], int neq, fvh_forest **out);
void fvh_forest_destroy(fvh_forest *fr);
RUN *fvh_forest_run(fvh_forest *fr);          /* the RUN whose topo/con/par describe this forest */
int fvh_forest_nleaves(fvh_forest *fr);       /* returns the count (>=0) */
int fvh_forest_max_level(fvh_forest *fr);     /* deepest leaf level (roots = 1) */
/* The creation arguments back (any pointer may be NULL). */
int fvh_forest_geometry(fvh_forest *fr, int nroot[3], double x0[3], double x1[3], int bc[6],
                        int periodic[3], int *neq);
/* Per leaf (list order): centre xc[3N], level[N] (reference levels, roots = 1). */
int fvh_forest_leaf_info(fvh_forest *fr, double *xc, int32_t *level);
/* Copy S/SN/RHS/SF/SG between leaf arrays and var-major host buffers (layouts as FVG_FIELD_*). */
int fvh_forest_set_field(fvh_forest *fr, int which, const double *src);
int fvh_forest_get_field(fvh_forest *fr, int which, double *dst);
/* Adapt: flag[N] per leaf: +1 split, -1 merge request, 0 keep.  Splits copy the parent S
 * (split.c:492-502); a parent whose 8 kids are all leaves flagged -1 is merged with the mean
 * of the kids (merge.c:92-107).  2:1 face balance is restored by further splits.  Leaves
 * deeper than max_level are not split.  Returns the new leaf count or an error. */
int fvh_forest_adapt(fvh_forest *fr, const int8_t *flag, int max_level);

/* ---------------------------------------------------------------- restart / export */
/* Binary restart of this library (no counterpart in the reference): the forest's box,
 * roots, side types and periodic flags, neq, tstep and time, the leaf levels in list order
 * and S var-major in f64; the load rebuilds the same forest with the same S bits. */
int fvh_forest_save(fvh_forest *fr, const char *path, int32_t tstep, double time);
int fvh_forest_load(const char *path, fvh_forest **out, int32_t *tstep, double *time);
/* The reference's text restart: replaces restart_save (restart_save.c:3-45; header
 * "tstep time nleaves ntrees neq", per leaf "root tag" and S at %15.10e) and
 * restart_load (restart_load.c:19-63, leaves spawned along their tags as spawntree.c).
 * The load wants a forest of bare roots (level 0) of the same geometry; every leaf of the
 * result must be listed once.  Tags hold 10 levels at most (FVG_ERR_UNSUPPORTED). */
int fvh_forest_save_text(fvh_forest *fr, const char *path, int32_t tstep, double time);
int fvh_forest_load_text(fvh_forest *fr, const char *path, int32_t *tstep, double *time);
/* Tecplot FEPOINT brick zone of the leaves of a RUN: replaces exportfield_linked_V3
 * (exportfield_linked_V3.c:19-572, one rank).  Nodes are the reference's links (coincident
 * element nodes joined across same-level and coarse-to-fine faces) with member-mean values
 * of S, p, |grad rho|, |grad rho u|, |grad rho v|, |grad p|; S and SG come from the leaves,
 * p[ncell] from the caller (fvg_pressure).  model 1..3 picks the variable names (0: take
 * run->con->model). */
int fvh_export_tecplot(RUN *run, const char *path, int model, const double *p);

/* ---------------------------------------------------------------- flattener */
/* Flatten the leaf list of `run` (any RUN whose leaves carry the reference's neighbour and
 * geometry fields).  The result owns its arrays; fvh_flat_mesh points into them. */
int fvh_flatten_run(RUN *run, fvh_flat **out);
const fvg_mesh *fvh_flat_mesh(const fvh_flat *fl);
void fvh_flat_destroy(fvh_flat *fl);
/* The per-level tables of leaves lv[0..n) (list order) exactly as fvh_flatten_run derives
 * them, evaluated only at the leaves that first set each entry (owners, as
 * fvg_remesh_begin reports them; -1 for an entry no leaf sets).  h, dist, area, vol hold
 * 32 rows; *nlevel receives the rows used (deepest level + 2). */
int fvh_tables_from_owners(BRANCH *const *lv, int n, const int32_t *owners, double *h,
                           double *dist, double *area, double *vol, int32_t *nlevel);

/* ---------------------------------------------------------------- device context */
int fvg_create(int device, const fvg_params *p, fvg_ctx **out);
int fvg_destroy(fvg_ctx *ctx);
int fvg_set_params(fvg_ctx *ctx, const fvg_params *p);
int fvg_set_mesh(fvg_ctx *ctx, const fvg_mesh *m);
int fvg_ncell(const fvg_ctx *ctx);
/* LDS-tiled stages (extension, no counterpart in the reference): fvg_set_mesh finds the
 * interior tiles of the mesh (8 x 8 x 8 same-level blocks of consecutive cells whose six
 * neighbour blocks are such tiles too); the out-of-place stages of fvg_steps /
 * fvg_time_steps update their cells with one kernel that keeps the face fluxes in LDS,
 * with results bitwise equal to the face-list path's.  The face flux buffer
 * (FVG_FIELD_FACEFLUX) then holds only the faces outside them.  mode: FVG_TILING_OFF,
 * FVG_TILING_ON (whenever the mesh has interior tiles), FVG_TILING_AUTO (the default:
 * the ideal gas, plain FP64 or viscous, with at least 256 interior tiles holding half of
 * the leaves or more, where the tiled stage measured faster; the stiff models, the mixed
 * Euler form and meshes with few tiles keep the face-list kernels).  fvg_tile_count = the
 * interior tiles of the current mesh. */
enum { FVG_TILING_OFF = 0, FVG_TILING_ON = 1, FVG_TILING_AUTO = 2 };
int fvg_set_tiling(fvg_ctx *ctx, int mode);
/* The interior tiles fvg_set_mesh would find in m, host only: *ntile = their count, and
 * the first min(ntile, cap) rows [8] (first leaf; first leaves of the neighbour tiles
 * across faces 0..5; level) into rows. */
int fvh_tile_table(const fvg_mesh *m, int32_t *rows, int cap, int *ntile);
int fvg_tile_count(const fvg_ctx *ctx);
/* HIP graphs (extension): fvg_steps and fvg_time_steps run their steps six at a time as
 * one captured graph (six steps bring the rotating state buffers back to where they
 * started), recaptured whenever the mesh, parameters, dt or method change; the last step
 * of a call always runs eagerly.  Results are bitwise those of eager launches.  Off by
 * default: the stage kernels already run back to back from the stream's queue, and the
 * graphs measured 0-3% slower on 69K to 2M leaves. */
int fvg_set_graphs(fvg_ctx *ctx, int on);
int fvg_set_state(fvg_ctx *ctx, const double *S);     /* host [neq][ncell] */
int fvg_get_state(fvg_ctx *ctx, double *S);           /* host [neq][ncell] */
int fvg_set_field(fvg_ctx *ctx, int which, const double *src);  /* S, SN, RHS, SF, SG from host f64 */
int fvg_get_field(fvg_ctx *ctx, int which, double *dst);        /* to host f64 */
/* Device pointer of a field (stream-ordered on fvg_stream).  SF/FACEFLUX are f32 when mixed. */
int fvg_device_ptr(fvg_ctx *ctx, int which, void **ptr);
void *fvg_stream(fvg_ctx *ctx);                       /* the context's hipStream_t */

/* The reference's phases, each one stream-ordered kernel sequence. */
int fvg_faceinterp(fvg_ctx *ctx);            /* SF from S (faceinterp.c:21-61) */
int fvg_gradcalc(fvg_ctx *ctx);              /* SG[neq+1][3] from S (gradcalc.c:9-47) */
/* Refinement flags per cell (host int8 [ncell]: +1 split, -1 merge, 0 keep) from the density
 * gradient of S, computed on the device so that only the flags cross to the host (the
 * criterion of criterionsplt.c:60-69 before the host's meshadaptation): kind 0 uses
 * h |grad rho| / rho, kind 1 |grad rho|; split where it exceeds `refine` below max_level,
 * merge where it is under `coarsen` above base_level.  Leaves SG = gradcalc(S). */
int fvg_adapt_flags(fvg_ctx *ctx, int kind, double refine, double coarsen, int base_level,
                    int max_level, int8_t *flags);
/* The same flags without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving or AMR
 * loops, harnesses that forbid hipMalloc).  The flags are staged in a static device array
 * of FVG_STATIC_FLAGS_MAX bytes that each GPU the library runs on holds until the process
 * ends; meshes with more cells get FVG_ERR_ARG.  Every context on one GPU shares that
 * array, so calls that overlap on one GPU (several contexts, several host threads) race;
 * each call overwrites the first ncell bytes and reads back only those. */
#define FVG_STATIC_FLAGS_MAX (1 << 24)
int fvg_adapt_flags_static(fvg_ctx *ctx, int kind, double refine, double coarsen, int base_level,
                           int max_level, int8_t *flags);
/* Cell pressure (host f64 [ncell]) from the device S, with the FP64 EOS of the gradient
 * kernel: the P column of the export (volmvalues.c:22-40, exportfield_linked_V3.c:390). */
int fvg_pressure(fvg_ctx *ctx, double *p);

/* ---------------------------------------------------------------- device re-flatten
 * After a host adaptation the mesh is rebuilt on the device from the leaves' keys, and
 * the state carried over from the previous mesh by key (no counterpart in the reference,
 * whose meshadaptation rebuilds its lists on the host; this replaces fvh_flatten_run +
 * fvg_set_mesh + the state round trip through the host for forests of this library).
 * begin: ncell leaves in list order, level[i] (roots = 1) and ijk[3i..3i+2], the global
 * cell index at that level (roots form an nroot grid); periodic axes and the bc of the six
 * domain sides (-x,+x,-y,+y,-z,+z; 1 or 2).  Builds classes, neighbours and face counts on
 * the device and returns, in owners[FVG_REMESH_OWNERS], the leaves that first set each
 * per-level table entry (rows [0,32) h, [32,64) volume, [64,160) area per level and
 * axis; -1 if unset).  FVG_ERR_STATE if a face has no leaf within one level (2:1).
 * end: the tables (32-row arrays as fvh_tables_from_owners fills them, nlevel rows used)
 * complete the mesh, which replaces the current one.  carry_state != 0 fills S from the
 * previous mesh (which must have been built by this path): a leaf present before keeps its
 * S, a split leaf's kids take the parent's S, a merged parent the mean of its 8 kids. */
#define FVG_REMESH_OWNERS 160
int fvg_remesh_begin(fvg_ctx *ctx, int32_t ncell, const int32_t *level, const int32_t *ijk,
                     const int32_t nroot[3], const int32_t periodic[3], const int32_t bc[6],
                     int32_t *owners);
int fvg_remesh_end(fvg_ctx *ctx, int32_t nlevel, const double *h, const double *dist,
                   const double *area, const double *vol, int carry_state);
int fvg_has_keys(const fvg_ctx *ctx);  /* 1 if the current mesh was built by fvg_remesh_* */
/* The forest's leaves in list order (valid until the next adaptation). */
BRANCH *const *fvh_forest_leaves(fvh_forest *fr);
/* Both steps for a forest of this library: keys from its leaf list, tables from its leaves. */
int fvh_forest_remesh_device(fvh_forest *fr, fvg_ctx *ctx, int carry_state);
/* The device mesh back on the host (for comparison with a host-flattened mesh; no
 * counterpart in the reference).  sizes[12] = ncell, nface, nface_axis[3], nquad, nlevel,
 * nface_reg[3], has_keys, 0.  Any array may be null; lengths as in fvg_mesh, tab holds
 * 34 * nlevel doubles (h, dist, area, vol, 1/d_cc). */
int fvg_get_mesh(fvg_ctx *ctx, int32_t *sizes, uint32_t *cflag, int32_t *nbr, int32_t *quad_nbr,
                 int32_t *face_L, int32_t *face_R, uint32_t *face_meta, int32_t *face_LL,
                 int32_t *face_RR, int32_t *gslot, double *tab);
int fvg_rhs(fvg_ctx *ctx);                   /* RHS from S / SF / SG (RHS.c:3-13) */
int fvg_rk(fvg_ctx *ctx, int irk, double dt);/* stage formula irk = 0..4 (RK.c:16-51) */
/* One time step: RK(0), then per stage [faceinterp] [gradcalc] RHS RK(k) with the RHS
 * gather fused into the stage update.  method 1: Euler, 2: SSP-RK2, 3: SSP-RK3. */
int fvg_step(fvg_ctx *ctx, int method, double dt);
int fvg_steps(fvg_ctx *ctx, int nsteps, int method, double dt);
/* dt = cfl / max_i sum_a (|u_a| + c_i) / h_a over the current S (extension; the reference
 * has a fixed dt). */
int fvg_cfl_dt(fvg_ctx *ctx, double cfl, double *dt);
/* out[0..neq-1] = sum_i S_i Vol_i, out[neq] = sum_i Vol_i (conservationcalc.c:31-38),
 * fixed-order device reduction. */
int fvg_totals(fvg_ctx *ctx, double *out);
/* Scan S and RHS for NaN/Inf or rho <= 0; *bad_cell = first offending cell or -1. */
int fvg_check_state(fvg_ctx *ctx, int32_t *bad_cell);
int fvg_sync(fvg_ctx *ctx);
const char *fvg_last_error(const fvg_ctx *ctx);
/* nsteps time steps; *ms_total = their device time (two events around them).  Then
 * ms_kernel[k] = the mean duration (ms) of each kernel of a stage, by kernel id (1
 * gradient, 2 face flux, 3 gather+RK, 4 LDS-tiled stage, 5 LDS-tiled gradient; 0 for a
 * kernel the stage does not launch), from batches of back-to-back launches of the same
 * work after the steps (the state is left as the steps made it). */
int fvg_time_steps(fvg_ctx *ctx, int nsteps, int method, double dt, double *ms_total,
                   double *ms_kernel /* [8] */);

/* ---------------------------------------------------------------- adjoint / tangent */
/* Rollout objective J = sum_{n=1..nsteps} dt sum_{i in probe} p(S^n_i) Vol_i of nsteps SSP-RK3
 * steps at a fixed dt from the context's S (model 1, inviscid; either limiter).  This is an
 * extension: the reference has no adjoint. */
int fvg_adj_set_probe(fvg_ctx *ctx, const uint8_t *mask /* host [ncell], 1 = in the probe box */);
/* Runs the rollout, keeping S every `ckpt` steps; *J = objective; S ends at the final state.
 * ckpt = 0 picks the widest spacing whose stores fit in half the free device memory: the
 * whole rollout where it fits, so that the reverse sweep recomputes nothing.  The spacing
 * does not change the gradient. */
int fvg_adj_forward(fvg_ctx *ctx, int nsteps, int ckpt, double dt, double *J);
/* The checkpoint spacing of the last fvg_adj_forward (0 before one). */
int fvg_adj_checkpoint_spacing(const fvg_ctx *ctx);
/* The objective of the last fvg_adj_forward / fvg_tlm as an unevaluated double-double sum
 * J_hi_lo[0] + J_hi_lo[1] (accumulated with error-free transformations), for finite
 * differences finer than the rounding of J itself. */
int fvg_adj_objective(const fvg_ctx *ctx, double *J_hi_lo /* [2] */);
/* Discrete adjoint dJ/dS^0 [neq][ncell] of the last forward rollout, recomputing stage
 * states from the checkpoints; into host memory and/or device memory (either may be NULL). */
int fvg_adj_backward(fvg_ctx *ctx, double *grad_host, void *grad_dev);
/* Tangent-linear rollout from the context's S along dir [neq][ncell] (host): *J and the
 * directional derivative *dJ (forward-mode, FP64 only).  S ends at the final state. */
int fvg_tlm(fvg_ctx *ctx, const double *dir, int nsteps, double dt, double *J, double *dJ);
/* Host evaluations of the device face-flux code for derivative checks (model 1, inviscid):
 * flux (F.n + HLLC correction, before -Area) across a face of normal sign*e_axis, its
 * vector-Jacobian product (reverse) and Jacobian-vector product (dual numbers), and the
 * limited slope with its partials w.r.t. delta0 and delta1. */
int fvg_face_flux_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                       double *out);
int fvg_face_vjp_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                      const double *ob, double *ULb, double *URb);
/* The same with the viscous flux (model 1, ns): gL / gR are each side's 9 momentum gradients
 * (d(rho u_i)/dx_j at [3 i + j]); the VJP also returns gb, the bar of either side's gradients
 * (the face uses their mean, so both sides get the same bar). */
int fvg_face_flux_ns_host(const fvg_params *p, int axis, int sign, const double *UL,
                          const double *UR, const double *gL, const double *gR, double *out);
int fvg_face_vjp_ns_host(const fvg_params *p, int axis, int sign, const double *UL,
                         const double *UR, const double *gL, const double *gR, const double *ob,
                         double *ULb, double *URb, double *gb);
int fvg_face_jvp_host(const fvg_params *p, int axis, int sign, const double *UL, const double *UR,
                      const double *ULd, const double *URd, double *outd);
/* Host evaluation of the device equation of state (eos64: ideal / stiffened gas, the
 * barotropic liquid-vapour law with its series power near rho_sat, the two-phase root) of
 * one conserved vector: out[0] = p, out[1] = c.  FVG_ERR_NONFINITE on a failed root. */
int fvg_eos_host(const fvg_params *p, const double *U, double *out /* [2] */);
int fvg_limiter_host(int limiter, double eps, double d0, double d1, double dp1cc, double dm1cc,
                     double *g, double *g0, double *g1);
/* Both face states' partials of one axis (the adjoint's reconstruction transpose):
 * out = (g0, g1) of SF[fp] = g(dP, dM; ip, im) and (h0, h1) of SF[fm] = g(dM, dP; im, ip),
 * zero on a boundary face (bp, bm).  Host check of the shared van Albada evaluation; no
 * counterpart in the reference. */
int fvg_limiter_axis_host(int limiter, double eps, int bp, int bm, double dP, double dM,
                          double ip, double im, double *out);
/* Bandwidth probe (no counterpart in the reference): 1 + reps copies of `bytes` (a multiple
 * of 16, 16-byte aligned device pointers on `device`) by a streaming kernel of 16-byte
 * loads and stores on a stream of its own; *ms = the mean duration of one of the last
 * reps copies (HIP events around them).  It gives the HBM rate a plain copy reaches,
 * beside the 8 TB/s peak, and a kernel of known traffic for the rocprof byte counters. */
int fvg_stream_copy(int device, void *dst, const void *src, size_t bytes, int reps, double *ms);
/* device-to-device state copies on the context's stream (torch interop) */
int fvg_set_state_device(fvg_ctx *ctx, const void *dS /* device [neq][ncell] f64 */);
int fvg_get_state_device(fvg_ctx *ctx, void *dS);

/* ---------------------------------------------------------------- RUN binding */
/* Flatten `run`, create a context on `device` and upload S; afterwards RHS(run), RK(irk,run),
 * faceinterp(run) and gradcalc(run) run on the GPU.  compat != 0 copies every output back
 * into the leaf arrays after each call (reference-shaped callers); compat == 0 keeps the
 * state on the device until fv_gpu_download(run).  `run` may come from the reference's
 * host code (its structs have the reference's layout, fv_strdata.h): the binding keeps
 * its state in run->info, a field the reference declares and never uses; attach
 * overwrites it without reading it, detach clears it, and the four entry points on a RUN
 * whose info is NULL do nothing but report FVG_ERR_STATE. */
int fv_gpu_attach(RUN *run, int device, const fvg_params *p, int compat);
int fv_gpu_detach(RUN *run);
int fv_gpu_remesh(RUN *run);      /* after a mesh adaptation: re-flatten and upload S */
int fv_gpu_upload(RUN *run);      /* leaves -> device S */
int fv_gpu_download(RUN *run);    /* device S, SN, RHS -> leaves */
fvg_ctx *fv_gpu_context(RUN *run);
int fv_gpu_last_status(RUN *run); /* status of the last RHS/RK/faceinterp/gradcalc call on run */

#ifdef __cplusplus
}
#endif

#endif /* FVGPU_H */
