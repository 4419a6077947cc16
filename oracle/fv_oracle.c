/*
 * fv_oracle.c — CPU oracle for the RHS + RK hot path (TEST INFRASTRUCTURE).
 *
 * A plain-C, FP64 restatement of the reference's algorithm, walking the leaf
 * list of a RUN through the reference's own pointers exactly as the reference
 * does: every interior face is evaluated from both of its sides
 * (residual_faces_AMR.c:37-177), the MUSCL face states come from
 * faceinterp.c:21-61, the gradients from gradcalc.c:9-47, the stage update
 * from RK.c:16-51.  It is the checker of the GPU path: only tests/, the smoke
 * check and bench.py's CPU baseline may load it.  The product never calls it.
 *
 * The reference itself cannot be compiled here (strdata.h:12-21 includes
 * metis.h, parmetis.h and PETSc headers that are not installed), so this file
 * is pinned by what the mathematics fixes (tests/test_oracle_pins.py):
 *   flux_adv / HLLC        consistency F(U,U) = F(U)·n, the four-region choice,
 *                          Toro's integral consistency of both star states (every
 *                          row, S* != u), brute force 2-leaf sums;
 *   EOS                    closed forms of the ideal gas and barotropic_eos,
 *                          the quadratic root of the two-phase mixture;
 *   minmod / van Albada    closed forms on 3-point data incl. the tie, the 1e-10
 *                          gate and the boundary 0/0 fall-through;
 *   distance               closed forms (h/4,h/2), (h/4,3h/8), (h/8,3h/8);
 *   2:1 hanging mean       exact linear fields on the coarse side of every hanging
 *                          face (mean of the 4 finer leaves at their centroid);
 *   inlet ghost (bc 1)     a uniform inlet state is steady, supersonic inflow passes
 *                          F(U_inlet), the MUSCL ghost enters the limiter by its sign;
 *   RHS                    free-stream preservation on uniform and 2:1 meshes,
 *                          discrete conservation, mirror symmetry;
 *   gradcalc               exact gradients of linear fields (interior cells);
 *   RK                     the SSP-RK3 coefficients on a linear ODE;
 *   flux_viscous           the Newtonian stress of a linear shear field.
 * Parity unpinned (no check above reaches it): nothing in the ordering of
 * advancetimeexplicit (undefined in the reference, strdata.h:568) beyond the
 * natural reading RK(0), [faceinterp] [gradcalc] RHS RK(k) per stage.
 *
 * Deliberate deviations (each also present in the GPU path):
 *   - model 2 MUSCL evaluates the EOS of the right state (facercnstr.c:37 always
 *     writes the left slot, leaving the right p and c stale);
 *   - the inlet ghost of models 2/3 evaluates its own EOS (asinvalues.c:26 and
 *     asinvalues_multiphase.c:22 write the left slot);
 *   - gradcalc zeroes SG[neq] before accumulating (gradcalc.c:17-21 does not);
 *   - periodic faces (new) measure distance() across the periodic image;
 *   - van Albada limiter and the CFL time step are new.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fvgpu.h"

#define ORC_OK 0
#define ORC_ERR_UNSUPPORTED (-5)
#define ORC_ERR_NONFINITE (-4)
#define ORC_ERR_ARG (-1)

/* face state scratch, the fields of SOL (strdata.h:360-402) the path uses */
typedef struct OSOL {
  double r, u, v, w, e, e_internal, p_hydro, c;
  double vec[8];
  double wec[8];  /* primitive variables (iprimtv option) */
  double ymass;
  double ux, uy, uz, vx, vy, vz, wx, wy, wz;
  double tx, ty, tz; /* gradient of p / rho (heat conduction option) */
} OSOL;

typedef struct OCTX {
  const fvg_params *p;
  int neq;
  OSOL sol, sol1;
  double flux_adv[8][3];
  double flux_viscous[8][3];
  double flux_HLLC[8];
  double Us_temp[8];
  double delta0[8], delta1[8];
  const double *period;
  int bad_cell;
  int status;
} OCTX;

static int g_last_bad = -1;
int orc_last_bad_cell(void) { return g_last_bad; }

/* Threads for the per-leaf loops (OpenMP; 1 = serial).  Every leaf writes only
 * its own fields and reads its neighbours' inputs, so the results do not depend
 * on the thread count; an error reports the lowest failing leaf index. */
static int g_threads = 1;
int orc_set_threads(int n) {
  g_threads = n > 0 ? n : 1;
  return g_threads;
}

static BRANCH **leaf_array(RUN *run, int *n) {
  int cnt = 0;
  for (BRANCH *c = run->topo->locl; c; c = c->lnxt) ++cnt;
  BRANCH **L = (BRANCH **)malloc((size_t)(cnt ? cnt : 1) * sizeof(BRANCH *));
  if (!L) { *n = -1; return NULL; }
  int q = 0;
  for (BRANCH *c = run->topo->locl; c; c = c->lnxt) L[q++] = c;
  *n = cnt;
  return L;
}

/* first (lowest-index) failure of a parallel leaf loop */
static void note_failure(int rc, int leaf, int *err, int *bad) {
#pragma omp critical(orc_fail)
  {
    if (!*err || leaf < *bad) { *err = rc; *bad = leaf; }
  }
}

static double sq(double x) { return x * x; }

/* inputset.c:130-222 */
static int orc_neq(int model) { return model == 2 ? 4 : ((model == 1 || model == 3) ? 5 : -1); }

/* ------------------------------------------------------------------ EOS */

/* barotropic_eos.c:3-30 (constants parametrised, defaults = the reference's) */
static void orc_barotropic_eos(const fvg_params *p, OSOL *s) {
  const double nexp = p->baro_nexp, beta = p->baro_beta, cbar = p->baro_cbar;
  const double psat = p->baro_psat, rsat = p->baro_rsat;
  if (s->r >= rsat) {
    s->p_hydro = psat + beta * (pow((s->r / rsat), nexp) - 1.0);
    s->c = sqrt(beta * nexp * (pow((s->r / rsat), (nexp - 1))) / rsat);
  }
  if (s->r < rsat) {
    s->p_hydro = psat + cbar * ((1.0 / rsat) - (1.0 / s->r));
    s->c = sqrt(cbar / ((s->r) * (s->r)));
  }
}

/* barotropic_eos.c:32-129: larger root of p^2 + b p + d = 0, retry with c = c_mix on NaN */
static int orc_barotropic_multiphase_eos(const fvg_params *p, OSOL *s) {
  const double R_G = p->mix_RG, T_ref = p->mix_Tref, r_sat = p->mix_rsat, p_sat = p->mix_psat;
  double c_temp = p->mix_cliq;
  double b = sq(c_temp) * r_sat - p_sat - s->r * s->ymass * R_G * T_ref -
             s->r * sq(c_temp) * (1.0 - s->ymass);
  double d = (p_sat - sq(c_temp) * r_sat) * s->r * s->ymass * R_G * T_ref;
  double chk = sqrt(sq(b) - 4.0 * d);
  if (isnan(chk)) {
    c_temp = p->mix_cmix;
    b = sq(c_temp) * r_sat - p_sat - s->r * s->ymass * R_G * T_ref -
        s->r * sq(c_temp) * (1.0 - s->ymass);
    d = (p_sat - sq(c_temp) * r_sat) * s->r * s->ymass * R_G * T_ref;
    chk = sqrt(sq(b) - 4.0 * d);
    if (isnan(chk)) return ORC_ERR_NONFINITE;
  }
  double p1 = (-b + sqrt(sq(b) - 4.0 * d)) / 2.0;
  double p2 = (-b - sqrt(sq(b) - 4.0 * d)) / 2.0;
  s->p_hydro = p1 > p2 ? p1 : p2;
  s->c = c_temp;
  return ORC_OK;
}

/* state from a conserved vector: the body shared by facevalues.c:25-54,
 * facercnstr.c:20-59 (clip = mass-fraction clip of facercnstr.c:43-44) */
static int orc_fill(OCTX *o, OSOL *s, const double *U, int clip) {
  const fvg_params *p = o->p;
  s->r = U[0];
  s->u = U[1] / s->r;
  s->v = U[2] / s->r;
  s->w = U[3] / s->r;
  s->vec[0] = s->r;
  s->vec[1] = U[1];
  s->vec[2] = U[2];
  s->vec[3] = U[3];
  if (p->model == 1) {
    s->e = U[4];
    s->vec[4] = U[4];
    s->e_internal = (s->e / s->r) - 0.5 * (sq(s->u) + sq(s->v) + sq(s->w));
    s->p_hydro = s->r * (p->gamma - 1.0) * s->e_internal - p->gamma * p->pinf;
    s->c = sqrt(p->gamma * (s->p_hydro + p->pinf) / s->r);
  } else if (p->model == 2) {
    orc_barotropic_eos(p, s);
  } else if (p->model == 3) {
    s->ymass = U[4] / s->r;
    s->vec[4] = U[4];
    if (clip) {
      if (s->ymass < (p->ymin / 1000.0)) s->ymass = (p->ymin / 1000.0);
      if (s->ymass > (1.0 - (p->ymin / 1000.0))) s->ymass = 1.0 - (p->ymin / 1000.0);
    }
    if (orc_barotropic_multiphase_eos(p, s)) return ORC_ERR_NONFINITE;
  } else {
    return ORC_ERR_UNSUPPORTED;
  }
  return ORC_OK;
}

static void orc_load_grads(OSOL *s, const LEAF *el) {
  s->ux = el->SG[1][0]; s->uy = el->SG[1][1]; s->uz = el->SG[1][2];
  s->vx = el->SG[2][0]; s->vy = el->SG[2][1]; s->vz = el->SG[2][2];
  s->wx = el->SG[3][0]; s->wy = el->SG[3][1]; s->wz = el->SG[3][2];
}

static int orc_theta_grad(OCTX *o, BRANCH *b, double g[3]);
static int orc_heat_on(const OCTX *o);

/* the gradient of p / rho of leaf b into s (heat conduction option) */
static int orc_load_theta(OCTX *o, BRANCH *b, OSOL *s) {
  if (!orc_heat_on(o)) return ORC_OK;
  double g[3];
  int rc = orc_theta_grad(o, b, g);
  s->tx = g[0];
  s->ty = g[1];
  s->tz = g[2];
  return rc;
}

/* facevalues.c:3-82 */
static int orc_facevalues(OCTX *o, BRANCH *crnt, int in, int ifc, int iside, OSOL *s) {
  BRANCH *b = iside == 0 ? crnt : crnt->neigtr[ifc][in];
  LEAF *el = b->el;
  int rc = orc_fill(o, s, el->S, 0);
  if (o->p->ns == 1) orc_load_grads(s, el);
  if (!rc) rc = orc_load_theta(o, b, s);
  return rc;
}

/* facercnstr.c:3-89 */
static int orc_facercnstr(OCTX *o, BRANCH *crnt, int in, int ifc, int iside, OSOL *s) {
  BRANCH *b = crnt;
  if (iside == 1) {
    b = crnt->neigtr[ifc][in];
    ifc = crnt->neigfc[ifc];
  }
  LEAF *el = b->el;
  int rc = orc_fill(o, s, el->SF[ifc], 1);
  if (o->p->ns == 1) orc_load_grads(s, el);
  if (!rc) rc = orc_load_theta(o, b, s);
  return rc;
}

/* asinvalues.c:3-66 / asinvalues_multiphase.c for the inlet ghost (bc_inlet.c) */
static int orc_bc(OCTX *o, int bc, OSOL *s1) {
  switch (bc) {
    case 0:
    case 2:
      return ORC_OK; /* bc.c:23-35: interior / outflow leave the copied state */
    case 1: {
      double U[8];
      for (int v = 0; v < o->neq; ++v) U[v] = o->p->inlet[v];
      return orc_fill(o, s1, U, 0);
    }
    default:
      return ORC_ERR_UNSUPPORTED; /* bc_reflective / bc_wall are undefined (bc.c:38,42) */
  }
}

/* cons2primtv.c:3-25 in its defined form (option iprimtv = 1): the density of the state
 * itself (the reference divides by a stale sol->r) and the kinetic energy from the
 * velocities (cons2primtv.c:17 squares the momenta) */
static void orc_cons2prim(const OCTX *o, const double *U, double *W) {
  const fvg_params *p = o->p;
  double r = U[0];
  W[0] = r;
  W[1] = U[1] / r;
  W[2] = U[2] / r;
  W[3] = U[3] / r;
  if (p->model == 1) {
    double ei = (U[4] / r) - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    W[4] = r * (p->gamma - 1.0) * ei - p->pinf;
  } else if (p->model == 3) {
    W[4] = U[4] / r;
  }
}

/* prim2conc.c:3-23 */
static void orc_prim2cons(const OCTX *o, const double *W, double *U) {
  const fvg_params *p = o->p;
  double r = W[0];
  U[0] = r;
  U[1] = r * W[1];
  U[2] = r * W[2];
  U[3] = r * W[3];
  if (p->model == 1) {
    double ei = (W[4] + p->pinf) / (r * (p->gamma - 1.0));
    U[4] = r * (ei + 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]));
  } else if (p->model == 3) {
    U[4] = r * W[4];
  }
}

/* Green-Gauss gradient of theta = p / rho of leaf b (option prandtl > 0): the loop of
 * gradcalc.c:23-40 over the faces and sub-faces of b, face value (theta_b + theta_nb)/2,
 * theta_nb of the neighbour's cell state or of the ghost (bc 2: b's own, bc 1: the inlet
 * state) */
static int orc_theta_grad(OCTX *o, BRANCH *b, double g[3]) {
  OSOL t;
  int rc = orc_fill(o, &t, b->el->S, 0);
  if (rc) return rc;
  const double th = t.p_hydro / t.r;
  const CELL *cl = b->cl;
  g[0] = g[1] = g[2] = 0.0;
  for (int f = 0; f < b->nlfc; ++f)
    for (int k = 0; k < b->nsfc[f]; ++k) {
      double thn = th;
      const int bc = cl->fc[f].bc;
      if (bc == 0) {
        rc = orc_fill(o, &t, b->neigtr[f][k]->el->S, 0);
        if (rc) return rc;
        thn = t.p_hydro / t.r;
      } else if (bc == 1) {
        double U[8];
        for (int v = 0; v < o->neq; ++v) U[v] = o->p->inlet[v];
        rc = orc_fill(o, &t, U, 0);
        if (rc) return rc;
        thn = t.p_hydro / t.r;
      }
      double m = 0.5 * (thn + th);
      g[0] += m * cl->nx[f] * cl->Area[f][k] / cl->Vol;
      g[1] += m * cl->ny[f] * cl->Area[f][k] / cl->Vol;
      g[2] += m * cl->nz[f] * cl->Area[f][k] / cl->Vol;
    }
  return ORC_OK;
}

static int orc_heat_on(const OCTX *o) { return o->p->model == 1 && o->p->ns == 1 && o->p->prandtl > 0.0; }

/* ------------------------------------------------------------------ MUSCL */

static void centre_of(const BRANCH *b, double c[3]) {
  double nd = (double)b->nlnd;
  c[0] = c[1] = c[2] = 0.0;
  for (int i = 0; i < b->nlnd; ++i) {
    c[0] += b->cl->nd[i].x / nd;
    c[1] += b->cl->nd[i].y / nd;
    c[2] += b->cl->nd[i].z / nd;
  }
}

static double orc_dist3(const OCTX *o, const double a[3], const double b[3]) {
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    double x = a[d] - b[d];
    double P = o->period ? o->period[d] : 0.0;
    if (P > 0.0) x -= P * nearbyint(x / P);
    s += sq(x);
  }
  return sqrt(s);
}

/* distance.c:3-148 */
static void orc_distance(const OCTX *o, BRANCH *crnt, int f, double *dcf, double *dcc) {
  double a[3], b[3], t;
  *dcf = 0.0;
  *dcc = 0.0;
  BRANCH *nb = crnt->neigtr[f][0];
  if (nb->level - crnt->level == 0) {
    centre_of(crnt, a);
    centre_of(nb, b);
    t = 0.5 * orc_dist3(o, a, b);
    *dcf = 0.5 * t;
    *dcc = t;
  }
  if (nb->level - crnt->level == 1) {
    centre_of(crnt, a);
    centre_of(nb->prnt, b);
    t = 0.5 * orc_dist3(o, a, b);
    *dcf = 0.5 * t;
    *dcc = 0.5 * t + 0.25 * t;
  }
  if ((nb->level - crnt->level == -1) && (nb->level != 0)) {
    centre_of(crnt->prnt, a);
    centre_of(nb, b);
    t = 0.25 * orc_dist3(o, a, b);
    *dcf = 0.25 * t;
    *dcc = 0.25 * t + 0.5 * t;
  }
}

/* deltavalues.c:3-26 with facevalues_MUSCL.c:3-56 */
static int orc_deltavalues(OCTX *o, BRANCH *crnt, double *delta, int f) {
  const int neq = o->neq;
  OSOL *s = &o->sol, *s1 = &o->sol1;
  for (int v = 0; v < neq; ++v) s->vec[v] = crnt->el->S[v];
  if (crnt->cl->fc[f].bc == 0) {
    if (crnt->lfc_neigs[f] == 1) {
      for (int v = 0; v < neq; ++v) s1->vec[v] = crnt->neigtr[f][0]->el->S[v];
    } else {
      int nn = crnt->lfc_neigs[f];
      for (int v = 0; v < neq; ++v) s1->vec[v] = 0.0;
      for (int k = 0; k < nn; ++k)
        for (int v = 0; v < neq; ++v) s1->vec[v] += crnt->neigtr[f][k]->el->S[v] / ((double)nn);
    }
  } else {
    for (int v = 0; v < neq; ++v) s1->vec[v] = crnt->el->S[v];
    if (crnt->cl->fc[f].bc == 1) {
      double U[8];
      for (int v = 0; v < neq; ++v) U[v] = o->p->inlet[v];
      int rc = orc_fill(o, s1, U, 0);
      if (rc) return rc;
    } else if (crnt->cl->fc[f].bc != 2) {
      return ORC_ERR_UNSUPPORTED;
    }
  }
  if (o->p->iprimtv == 1) {
    /* deltavalues.c:16-19: the deltas of the primitive variables */
    orc_cons2prim(o, s->vec, s->wec);
    orc_cons2prim(o, s1->vec, s1->wec);
    for (int v = 0; v < neq; ++v) delta[v] = s1->wec[v] - s->wec[v];
    return ORC_OK;
  }
  for (int v = 0; v < neq; ++v) delta[v] = s1->vec[v] - s->vec[v];
  return ORC_OK;
}

/* reconctruct_V3.c:3-57; limiter FVG_VANALBADA is an extension.  iprimtv = 1: the
 * primitive face value (reconctruct_V3.c:37-38) through prim2conc (:47-51) */
static void orc_reconstruct(OCTX *o, BRANCH *crnt, int f, double dp1cf, double dp1cc,
                            double dm1cf, double dm1cc) {
  (void)dm1cf;
  const int prim = o->p->iprimtv == 1;
  double Wf[8];
  for (int v = 0; v < o->neq; ++v) {
    double grad = 0.0;
    if (crnt->cl->fc[f].bc == 0) {
      if (o->p->limiter == FVG_MINMOD) {
        if (fabs(o->delta0[v]) > 1e-10) {
          double prod = (-o->delta1[v] / dm1cc) * (o->delta0[v] / dp1cc);
          if (prod > 0.0) {
            if (fabs(-o->delta1[v] / dm1cc) > fabs(o->delta0[v] / dp1cc)) grad = o->delta0[v] / dp1cc;
            else grad = -o->delta1[v] / dm1cc;
          } else {
            grad = 0.0;
          }
        }
      } else {
        /* smooth van Albada of a = backward and b = forward slope; a zero
         * distance (boundary behind) contributes a zero slope */
        double a = dm1cc != 0.0 ? -o->delta1[v] / dm1cc : 0.0;
        double b = dp1cc != 0.0 ? o->delta0[v] / dp1cc : 0.0;
        double e = o->p->vanalbada_eps;
        grad = (a * (b * b + e) + b * (a * a + e)) / (a * a + b * b + 2.0 * e);
      }
    }
    if (prim) Wf[v] = o->sol.wec[v] + dp1cf * grad;
    else crnt->el->SF[f][v] = o->sol.vec[v] + dp1cf * grad;
  }
  if (prim) orc_prim2cons(o, Wf, crnt->el->SF[f]);
}

static const int OFCOPP[6] = {2, 3, 0, 1, 5, 4};

/* faceinterp.c:21-61; period[a] > 0 is the length of a periodic axis (an extension: the
 * distance to a neighbour across the periodic side is measured to its image), NULL or 0
 * for none */
int orc_faceinterp(RUN *run, const fvg_params *p, const double *period) {
  int n, err = 0, bad = -1;
  BRANCH **L = leaf_array(run, &n);
  if (!L) return ORC_ERR_ARG;
#pragma omp parallel num_threads(g_threads)
  {
    OCTX o;
    memset(&o, 0, sizeof(o));
    o.p = p;
    o.neq = run->con->neq;
    o.period = period;
#pragma omp for schedule(static)
    for (int q = 0; q < n; ++q) {
      BRANCH *c = L[q];
      for (int f = 0; f < c->nlfc; ++f) {
        int fo = OFCOPP[f];
        double p1cf, p1cc, m1cf, m1cc;
        orc_distance(&o, c, f, &p1cf, &p1cc);
        orc_distance(&o, c, fo, &m1cf, &m1cc);
        int rc = orc_deltavalues(&o, c, o.delta1, fo);
        if (!rc) rc = orc_deltavalues(&o, c, o.delta0, f);
        if (rc) { note_failure(rc, q, &err, &bad); break; }
        orc_reconstruct(&o, c, f, p1cf, p1cc, m1cf, m1cc);
      }
    }
  }
  free(L);
  if (err) { g_last_bad = bad; return err; }
  return ORC_OK;
}

/* ------------------------------------------------------------------ gradients */

/* gradcalc.c:9-47 */
int orc_gradcalc(RUN *run, const fvg_params *p) {
  fvg_params q = *p;
  q.ns = 0; /* face states here need no gradients */
  const int neq = run->con->neq;
  int n, err = 0, bad = -1;
  BRANCH **Lv = leaf_array(run, &n);
  if (!Lv) return ORC_ERR_ARG;
#pragma omp parallel num_threads(g_threads)
  {
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = &q;
  o.neq = neq;
#pragma omp for schedule(static)
  for (int iq = 0; iq < n; ++iq) {
    BRANCH *b = Lv[iq];
    int failed = 0;
    for (int v = 0; v <= neq; ++v) b->el->SG[v][0] = b->el->SG[v][1] = b->el->SG[v][2] = 0.0;
    for (int f = 0; f < b->nlfc; ++f) {
      for (int k = 0; k < b->nsfc[f]; ++k) {
        int rc = orc_facevalues(&o, b, k, f, 0, &o.sol);
        if (!rc) {
          if (b->cl->fc[f].bc == 0) rc = orc_facevalues(&o, b, k, f, 1, &o.sol1);
          else {
            rc = orc_facevalues(&o, b, k, f, 0, &o.sol1);
            if (!rc) rc = orc_bc(&o, b->cl->fc[f].bc, &o.sol1);
          }
        }
        if (rc) { note_failure(rc, iq, &err, &bad); failed = 1; break; }
        CELL *cl = b->cl;
        for (int v = 0; v < neq; ++v) {
          double m = 0.5 * (o.sol1.vec[v] + o.sol.vec[v]);
          b->el->SG[v][0] += m * cl->nx[f] * cl->Area[f][k] / cl->Vol;
          b->el->SG[v][1] += m * cl->ny[f] * cl->Area[f][k] / cl->Vol;
          b->el->SG[v][2] += m * cl->nz[f] * cl->Area[f][k] / cl->Vol;
        }
        double mp = 0.5 * (o.sol1.p_hydro + o.sol.p_hydro);
        b->el->SG[neq][0] += mp * cl->nx[f] * cl->Area[f][k] / cl->Vol;
        b->el->SG[neq][1] += mp * cl->ny[f] * cl->Area[f][k] / cl->Vol;
        b->el->SG[neq][2] += mp * cl->nz[f] * cl->Area[f][k] / cl->Vol;
      }
      if (failed) break;
    }
  }
  }
  free(Lv);
  if (err) { g_last_bad = bad; return err; }
  return ORC_OK;
}

/* ------------------------------------------------------------------ fluxes */

/* flux_adv.c:10-45 */
static void orc_flux_adv(OCTX *o, const OSOL *s) {
  double(*F)[3] = o->flux_adv;
  const double p = s->p_hydro;
  F[0][0] = s->u * s->r;
  F[0][1] = s->v * s->r;
  F[0][2] = s->w * s->r;
  F[1][0] = s->r * s->u * s->u + p;
  F[1][1] = s->r * s->u * s->v + 0.0;
  F[1][2] = s->r * s->u * s->w + 0.0;
  F[2][0] = s->r * s->v * s->u + 0.0;
  F[2][1] = s->r * s->v * s->v + p;
  F[2][2] = s->r * s->v * s->w + 0.0;
  F[3][0] = s->r * s->w * s->u + 0.0;
  F[3][1] = s->r * s->w * s->v + 0.0;
  F[3][2] = s->r * s->w * s->w + p;
  if (o->p->model == 1) {
    F[4][0] = s->u * (s->e + p);
    F[4][1] = s->v * (s->e + p);
    F[4][2] = s->w * (s->e + p);
  }
  if (o->p->model == 3) {
    F[4][0] = s->u * s->r * s->ymass;
    F[4][1] = s->v * s->r * s->ymass;
    F[4][2] = s->w * s->r * s->ymass;
  }
}

/* rotate_vectors.c:32-91 applied to st = p I; returns element [0][0] */
static double orc_rotate_s11(double p, const double R[3][3]) {
  double st[3][3] = {{p, 0.0, 0.0}, {0.0, p, 0.0}, {0.0, 0.0, p}};
  double M1[3][3], M2[3][3], RT[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { RT[j][i] = R[i][j]; M1[i][j] = 0.0; M2[i][j] = 0.0; }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) M1[i][j] += R[i][k] * st[k][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) M2[i][j] += M1[i][k] * RT[k][j];
  return M2[0][0];
}

/* star_region_euler.c:3-40 */
static void orc_star(OCTX *o, const OSOL *s, const double R[3][3], double SS, double SK,
                     double uK, double vK, double wK) {
  double cs = s->r * (SK - uK) / (SK - SS);
  double u = SS, v = vK, w = wK;
  o->Us_temp[0] = cs;
  o->Us_temp[1] = cs * (u * R[0][0] + v * R[1][0] + w * R[2][0]);
  o->Us_temp[2] = cs * (u * R[0][1] + v * R[1][1] + w * R[2][1]);
  o->Us_temp[3] = cs * (u * R[0][2] + v * R[1][2] + w * R[2][2]);
  if (o->p->model == 1)
    o->Us_temp[4] = cs * ((s->e / s->r) + (SS - uK) * (SS + (s->p_hydro / (s->r * (SK - uK)))));
  if (o->p->model == 3) o->Us_temp[4] = cs * s->ymass;
}

/* rotated velocities and wave speeds of the face states sol / sol1
 * (compute_wave_speeds.c:33-87: Davis S_L, S_R and the contact speed S*) */
typedef struct OWAVES {
  double uL, vL, wL, uR, vR, wR, SL, SS, SR;
} OWAVES;

static int orc_waves(const OCTX *o, const double R[3][3], OWAVES *w) {
  const OSOL *L = &o->sol, *Rs = &o->sol1;
  w->uL = L->u * R[0][0] + L->v * R[0][1] + L->w * R[0][2];
  w->vL = L->u * R[1][0] + L->v * R[1][1] + L->w * R[1][2];
  w->wL = L->u * R[2][0] + L->v * R[2][1] + L->w * R[2][2];
  w->uR = Rs->u * R[0][0] + Rs->v * R[0][1] + Rs->w * R[0][2];
  w->vR = Rs->u * R[1][0] + Rs->v * R[1][1] + Rs->w * R[1][2];
  w->wR = Rs->u * R[2][0] + Rs->v * R[2][1] + Rs->w * R[2][2];
  const double uL = w->uL, uR = w->uR;
  double ruL = L->r * uL, ruR = Rs->r * uR;
  double cL = L->c, cR = Rs->c;
  double SR = (uL + cL) > (uR + cR) ? (uL + cL) : (uR + cR);
  double SL = (uL - cL) < (uR - cR) ? (uL - cL) : (uR - cR);
  if (o->p->wave_speeds == FVG_WAVES_PVRS) {
    /* option: PVRS speeds (Toro, Riemann Solvers, 3rd ed., 10.5.2), pressures as p + p_inf */
    const double g = o->p->gamma, pL = L->p_hydro + o->p->pinf, pR = Rs->p_hydro + o->p->pinf;
    double ps = 0.5 * (pL + pR) - 0.5 * (uR - uL) * (0.5 * (L->r + Rs->r)) * (0.5 * (cL + cR));
    if (ps < 0.0) ps = 0.0;
    const double k = (g + 1.0) / (2.0 * g);
    const double qL = ps <= pL ? 1.0 : sqrt(1.0 + k * (ps / pL - 1.0));
    const double qR = ps <= pR ? 1.0 : sqrt(1.0 + k * (ps / pR - 1.0));
    SL = uL - cL * qL;
    SR = uR + cR * qR;
  }
  double s11L = orc_rotate_s11(L->p_hydro, R);
  double s11R = orc_rotate_s11(Rs->p_hydro, R);
  double SS = ((L->r * sq(uL) + s11L) - (Rs->r * sq(uR) + s11R) - (SL * ruL) + (SR * ruR)) /
              (ruL - ruR - (SL * L->r) + (SR * Rs->r));
  w->SL = SL;
  w->SS = SS;
  w->SR = SR;
  return isnan(SS) ? ORC_ERR_NONFINITE : ORC_OK;
}

/* hllc_euler.c:3-90 with compute_wave_speeds.c:3-90 (Davis speeds) */
static int orc_hllc(OCTX *o, BRANCH *crnt, int ifc) {
  const OSOL *L = &o->sol, *Rs = &o->sol1;
  const CELL *cl = crnt->cl;
  const double R[3][3] = {{cl->nx[ifc], cl->ny[ifc], cl->nz[ifc]},
                          {cl->nxt1[ifc], cl->nyt1[ifc], cl->nzt1[ifc]},
                          {cl->nxt2[ifc], cl->nyt2[ifc], cl->nzt2[ifc]}};
  OWAVES wv;
  if (orc_waves(o, R, &wv)) return ORC_ERR_NONFINITE;
  const double uL = wv.uL, vL = wv.vL, wL = wv.wL, uR = wv.uR, vR = wv.vR, wR = wv.wR;
  const double SL = wv.SL, SS = wv.SS, SR = wv.SR;
  const int neq = o->neq;
  for (int v = 0; v < neq; ++v) o->flux_HLLC[v] = 0.0;
  int flag = 0;
  if (0.0 <= SL && flag == 0) {
    flag = 1;
    orc_flux_adv(o, L);
  }
  if ((SL <= 0.0 && 0.0 <= SS) && flag == 0) {
    flag = 2;
    orc_star(o, L, R, SS, SL, uL, vL, wL);
    orc_flux_adv(o, L);
    for (int v = 0; v < neq; ++v) o->flux_HLLC[v] = SL * (o->Us_temp[v] - L->vec[v]);
  }
  if ((SS <= 0.0 && 0.0 <= SR) && flag == 0) {
    flag = 3;
    orc_star(o, Rs, R, SS, SR, uR, vR, wR);
    orc_flux_adv(o, Rs);
    for (int v = 0; v < neq; ++v) o->flux_HLLC[v] = SR * (o->Us_temp[v] - Rs->vec[v]);
  }
  if (SR <= 0.0 && flag == 0) {
    flag = 4;
    orc_flux_adv(o, Rs);
  }
  if (flag == 0) return ORC_ERR_NONFINITE;
  for (int v = 0; v < neq; ++v)
    if (isnan(o->flux_HLLC[v])) return ORC_ERR_NONFINITE;
  return ORC_OK;
}

/* flux_viscous.c:7-235 with turb = 0 (mu_turb = 0) */
static void orc_flux_viscous(OCTX *o) {
  const OSOL *s = &o->sol, *s1 = &o->sol1;
  double uf = 0.5 * (s->u + s1->u);
  double vf = 0.5 * (s->v + s1->v);
  double wf = 0.5 * (s->w + s1->w);
  double ux = 0.5 * (s->ux + s1->ux), uy = 0.5 * (s->uy + s1->uy), uz = 0.5 * (s->uz + s1->uz);
  double vx = 0.5 * (s->vx + s1->vx), vy = 0.5 * (s->vy + s1->vy), vz = 0.5 * (s->vz + s1->vz);
  double wx = 0.5 * (s->wx + s1->wx), wy = 0.5 * (s->wy + s1->wy), wz = 0.5 * (s->wz + s1->wz);
  double mu = o->p->mu + 0.0;
  double txx = mu * (2.0 * ux - (2.0 / 3.0) * (ux + vy + wz));
  double tyy = mu * (2.0 * vy - (2.0 / 3.0) * (ux + vy + wz));
  double tzz = mu * (2.0 * wz - (2.0 / 3.0) * (ux + vy + wz));
  double txy = mu * (uy + vx);
  double txz = mu * (uz + wx);
  double tyz = mu * (vz + wy);
  double tyx = txy, tzy = tyz, tzx = txz;
  double(*F)[3] = o->flux_viscous;
  F[0][0] = 0.0; F[0][1] = 0.0; F[0][2] = 0.0;
  F[1][0] = txx; F[1][1] = tyx; F[1][2] = tzx;
  F[2][0] = txy; F[2][1] = tyy; F[2][2] = tzy;
  F[3][0] = txz; F[3][1] = tyz; F[3][2] = tzz;
  if (o->p->visc_work && o->neq > 4) {
    F[4][0] = uf * txx + vf * txy + wf * txz;
    F[4][1] = uf * tyx + vf * tyy + wf * tyz;
    F[4][2] = uf * tzx + vf * tzy + wf * tzz;
  }
  if (orc_heat_on(o)) {
    /* option: -q = mu gamma / ((gamma - 1) Pr) grad(p / rho), face mean of the two cells' */
    const fvg_params *p = o->p;
    const double kh = p->mu * p->gamma / ((p->gamma - 1.0) * p->prandtl);
    const double w0 = p->visc_work ? F[4][0] : 0.0, w1 = p->visc_work ? F[4][1] : 0.0,
                 w2 = p->visc_work ? F[4][2] : 0.0;
    F[4][0] = w0 + kh * (0.5 * (s->tx + s1->tx));
    F[4][1] = w1 + kh * (0.5 * (s->ty + s1->ty));
    F[4][2] = w2 + kh * (0.5 * (s->tz + s1->tz));
  }
}

/* residual_faces_AMR.c:10-202 for one leaf */
static int orc_residual(OCTX *o, BRANCH *crnt) {
  const int neq = o->neq;
  const fvg_params *p = o->p;
  double flux_face[6][8];
  for (int v = 0; v < neq; ++v) {
    crnt->el->RHS[v] = 0.0;
    o->flux_viscous[v][0] = o->flux_viscous[v][1] = o->flux_viscous[v][2] = 0.0;
  }
  for (int f = 0; f < crnt->nlfc; ++f) {
    int Nneig;
    int bc = crnt->cl->fc[f].bc;
    if (bc == 0) Nneig = crnt->lfc_neigs[f] == 1 ? 1 : crnt->lfc_neigs[f];
    else Nneig = 1;
    for (int v = 0; v < neq; ++v) flux_face[f][v] = 0.0;
    int rc = p->ischeme == 1 ? orc_facevalues(o, crnt, 0, f, 0, &o->sol)
                             : orc_facercnstr(o, crnt, 0, f, 0, &o->sol);
    if (rc) return rc;
    for (int k = 0; k < Nneig; ++k) {
      if (bc == 0) {
        rc = p->ischeme == 1 ? orc_facevalues(o, crnt, k, f, 1, &o->sol1)
                             : orc_facercnstr(o, crnt, k, f, 1, &o->sol1);
      } else {
        rc = p->ischeme == 1 ? orc_facevalues(o, crnt, k, f, 0, &o->sol1)
                             : orc_facercnstr(o, crnt, k, f, 0, &o->sol1);
        if (!rc) rc = orc_bc(o, bc, &o->sol1);
      }
      if (rc) return rc;
      if (p->ns == 1) orc_flux_viscous(o);
      rc = orc_hllc(o, crnt, f);
      if (rc) return rc;
      const CELL *cl = crnt->cl;
      double A = cl->Area[f][k];
      for (int v = 0; v < neq; ++v) {
        flux_face[f][v] += -(A) * (((o->flux_adv[v][0] - o->flux_viscous[v][0]) * cl->nx[f] +
                                    (o->flux_adv[v][1] - o->flux_viscous[v][1]) * cl->ny[f] +
                                    (o->flux_adv[v][2] - o->flux_viscous[v][2]) * cl->nz[f]) +
                                   o->flux_HLLC[v]);
      }
    }
    for (int v = 0; v < neq; ++v) crnt->el->RHS[v] += flux_face[f][v];
  }
  for (int v = 0; v < neq; ++v)
    if (isnan(crnt->el->RHS[v])) return ORC_ERR_NONFINITE;
  return ORC_OK;
}

static int check_params(RUN *run, const fvg_params *p) {
  if (!run || !p || !run->con || !run->topo) return ORC_ERR_ARG;
  if (orc_neq(p->model) != run->con->neq) return ORC_ERR_ARG;
  if (p->ischeme != 1 && p->ischeme != 2) return ORC_ERR_UNSUPPORTED;
  return ORC_OK;
}

/* RHS.c:3-13 */
int orc_rhs(RUN *run, const fvg_params *p) {
  int rc = check_params(run, p);
  if (rc) return rc;
  int n, err = 0, bad = -1;
  BRANCH **L = leaf_array(run, &n);
  if (!L) return ORC_ERR_ARG;
#pragma omp parallel num_threads(g_threads)
  {
    OCTX o;
    memset(&o, 0, sizeof(o));
    o.p = p;
    o.neq = run->con->neq;
#pragma omp for schedule(static)
    for (int q = 0; q < n; ++q) {
      int r = orc_residual(&o, L[q]);
      if (r) note_failure(r, q, &err, &bad);
    }
  }
  free(L);
  if (err) { g_last_bad = bad; return err; }
  return ORC_OK;
}

/* RK.c:3-56 */
int orc_rk(RUN *run, int irk, double dt) {
  const int neq = run->con->neq;
  int n;
  BRANCH **Lv = leaf_array(run, &n);
  if (!Lv) return ORC_ERR_ARG;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int q = 0; q < n; ++q) {
    BRANCH *c = Lv[q];
    LEAF *el = c->el;
    double V = c->cl->Vol;
    for (int f = 0; f < c->nlfc; ++f) el->flux_flag[f] = 0;
    switch (irk) {
      case 0:
        for (int e = 0; e < neq; ++e) el->SN[e] = el->S[e];
        break;
      case 1:
        for (int e = 0; e < neq; ++e) el->S[e] = el->SN[e] + dt * (1.0 / V) * el->RHS[e];
        break;
      case 2:
        for (int e = 0; e < neq; ++e)
          el->S[e] = (3.0 / 4.0) * el->SN[e] + (1.0 / 4.0) * el->S[e] + (1.0 / 4.0) * dt * el->RHS[e] / V;
        break;
      case 3:
        for (int e = 0; e < neq; ++e)
          el->S[e] = (1.0 / 3.0) * el->SN[e] + (2.0 / 3.0) * el->S[e] + (2.0 / 3.0) * dt * el->RHS[e] / V;
        break;
      case 4:
        for (int e = 0; e < neq; ++e)
          el->S[e] = (1.0 / 2.0) * el->SN[e] + (1.0 / 2.0) * el->S[e] + (1.0 / 2.0) * dt * el->RHS[e] / V;
        break;
      default:
        break;
    }
  }
  free(Lv);
  return ORC_OK;
}

/* the natural reading of the undefined advancetimeexplicit (loop.c:48) */
int orc_step(RUN *run, const fvg_params *p, const double *period, int method, double dt) {
  static const int st3[3] = {1, 2, 3}, st2[2] = {1, 4}, st1[1] = {1};
  const int *st = method == 3 ? st3 : (method == 2 ? st2 : st1);
  int ns = method == 3 ? 3 : (method == 2 ? 2 : 1);
  int rc = orc_rk(run, 0, dt);
  for (int k = 0; k < ns && !rc; ++k) {
    if (p->ischeme == 2) rc = orc_faceinterp(run, p, period);
    if (!rc && p->ns == 1) rc = orc_gradcalc(run, p);
    if (!rc) rc = orc_rhs(run, p);
    if (!rc) rc = orc_rk(run, st[k], dt);
  }
  return rc;
}

/* conservationcalc.c:19-52 without the MPI reduction */
int orc_totals(RUN *run, double *out) {
  const int neq = run->con->neq;
  for (int v = 0; v <= neq; ++v) out[v] = 0.0;
  for (BRANCH *c = run->topo->locl; c; c = c->lnxt) {
    for (int v = 0; v < neq; ++v) out[v] += c->el->S[v] * c->cl->Vol;
    out[neq] += c->cl->Vol;
  }
  return ORC_OK;
}

/* CFL time step (extension): dt = cfl / max_i sum_a (|u_a| + c) / h_a */
int orc_cfl_dt(RUN *run, const fvg_params *p, double cfl, double *dt) {
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = p;
  o.neq = run->con->neq;
  double m = 0.0;
  int q = 0;
  for (BRANCH *c = run->topo->locl; c; c = c->lnxt, ++q) {
    int rc = orc_fill(&o, &o.sol, c->el->S, 0);
    if (rc) return rc;
    const ND *n = c->cl->nd;
    double hx = n[1].x - n[0].x, hy = n[2].y - n[0].y, hz = n[4].z - n[0].z;
    double r = (fabs(o.sol.u) + o.sol.c) / hx + (fabs(o.sol.v) + o.sol.c) / hy +
               (fabs(o.sol.w) + o.sol.c) / hz;
    if (!(r == r)) { g_last_bad = q; return ORC_ERR_NONFINITE; }
    if (r > m) m = r;
  }
  *dt = cfl / m;
  return ORC_OK;
}

/* single-face helpers used by the closed-form pins: HLLC flux of two conserved
 * states across a face with unit normal n and tangents t1, t2 (rows of R) */
int orc_face_flux(const fvg_params *p, const double *UL, const double *UR, const double *R9,
                  double *out) {
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = p;
  o.neq = orc_neq(p->model);
  int rc = orc_fill(&o, &o.sol, UL, 1);
  if (!rc) rc = orc_fill(&o, &o.sol1, UR, 1);
  if (rc) return rc;
  BRANCH b;
  CELL cl;
  double nx = R9[0], ny = R9[1], nz = R9[2], t1x = R9[3], t1y = R9[4], t1z = R9[5];
  double t2x = R9[6], t2y = R9[7], t2z = R9[8];
  cl.nx = &nx; cl.ny = &ny; cl.nz = &nz;
  cl.nxt1 = &t1x; cl.nyt1 = &t1y; cl.nzt1 = &t1z;
  cl.nxt2 = &t2x; cl.nyt2 = &t2y; cl.nzt2 = &t2z;
  b.cl = &cl;
  rc = orc_hllc(&o, &b, 0);
  if (rc) return rc;
  for (int v = 0; v < o.neq; ++v)
    out[v] = (o.flux_adv[v][0] * nx + o.flux_adv[v][1] * ny + o.flux_adv[v][2] * nz) + o.flux_HLLC[v];
  return ORC_OK;
}

/* Both star states of star_region_euler.c:3-40 for the face states UL | UR (whichever
 * region the face falls in), with the wave speeds (S_L, S*, S_R) they are built from;
 * for the integral-consistency pin of the star region. */
int orc_star_states(const fvg_params *p, const double *UL, const double *UR, const double *R9,
                    double *speeds, double *UsL, double *UsR) {
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = p;
  o.neq = orc_neq(p->model);
  int rc = orc_fill(&o, &o.sol, UL, 1);
  if (!rc) rc = orc_fill(&o, &o.sol1, UR, 1);
  if (rc) return rc;
  const double R[3][3] = {{R9[0], R9[1], R9[2]}, {R9[3], R9[4], R9[5]}, {R9[6], R9[7], R9[8]}};
  OWAVES wv;
  rc = orc_waves(&o, R, &wv);
  if (rc) return rc;
  speeds[0] = wv.SL;
  speeds[1] = wv.SS;
  speeds[2] = wv.SR;
  orc_star(&o, &o.sol, R, wv.SS, wv.SL, wv.uL, wv.vL, wv.wL);
  for (int v = 0; v < o.neq; ++v) UsL[v] = o.Us_temp[v];
  orc_star(&o, &o.sol1, R, wv.SS, wv.SR, wv.uR, wv.vR, wv.wR);
  for (int v = 0; v < o.neq; ++v) UsR[v] = o.Us_temp[v];
  return ORC_OK;
}

/* pressure and sound speed of one conserved state */
int orc_eos(const fvg_params *p, const double *U, double *pc) {
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = p;
  o.neq = orc_neq(p->model);
  int rc = orc_fill(&o, &o.sol, U, 0);
  pc[0] = o.sol.p_hydro;
  pc[1] = o.sol.c;
  return rc;
}

/* the limited slope of reconctruct_V3 for one variable */
double orc_limited_slope(int limiter, double eps, int boundary_face, double delta0, double delta1,
                         double dp1cc, double dm1cc) {
  fvg_params p;
  memset(&p, 0, sizeof(p));
  p.limiter = limiter;
  p.vanalbada_eps = eps;
  OCTX o;
  memset(&o, 0, sizeof(o));
  o.p = &p;
  o.neq = 1;
  o.delta0[0] = delta0;
  o.delta1[0] = delta1;
  double SFrow[1], S[1] = {0.0};
  double *SFp[6] = {SFrow, SFrow, SFrow, SFrow, SFrow, SFrow};
  LEAF el;
  FC fc[6];
  CELL cl;
  BRANCH b;
  memset(fc, 0, sizeof(fc));
  fc[0].bc = boundary_face ? 2 : 0;
  el.SF = SFp;
  el.S = S;
  cl.fc = fc;
  b.el = &el;
  b.cl = &cl;
  o.sol.vec[0] = 0.0;
  orc_reconstruct(&o, &b, 0, 1.0, dp1cc, 0.0, dm1cc);
  return SFrow[0];
}
