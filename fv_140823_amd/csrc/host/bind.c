/*
 * The reference's hot-path entry points with their own signatures
 * (strdata.h:642,654,656,675), bound to a device context:
 *   RHS(run)         -> fvg_rhs        (RHS.c:3-13)
 *   RK(irk, run)     -> fvg_rk         (RK.c:3-56), dt = run->con->dt
 *   faceinterp(run)  -> fvg_faceinterp (faceinterp.c:21-61)
 *   gradcalc(run)    -> fvg_gradcalc   (gradcalc.c:9-47)
 * fv_gpu_attach() flattens the leaf list once; after a mesh adaptation the
 * host calls fv_gpu_remesh().  In compat mode each call copies its outputs back
 * into the leaves (LEAF::RHS/S/SN/SF/SG), so reference-shaped callers work
 * unchanged; otherwise the state stays on the device until fv_gpu_download().
 *
 * The binding's state is per RUN, in run->info: the reference declares struct INFO
 * (strdata.h:423) but never defines it and no reference file reads or writes that
 * field, so this file defines it.  fv_gpu_attach() overwrites run->info without
 * reading it (a RUN on the reference's stack holds garbage there) and
 * fv_gpu_detach() clears it; the entry points on a RUN whose info is NULL report
 * FVG_ERR_STATE.  Only the leaf arrays the reference owns (S, SN, RHS, SF, SG) are
 * written; the leaves' structs are only read.
 * The reference reports errors by printf + exit(0) (residual_faces_AMR.c:186-194);
 * these wrappers keep the void signature and record the status in the binding
 * instead (fv_gpu_last_status(run)).
 */
#include <stdlib.h>
#include <string.h>

#include "fvgpu.h"

#define FV_BINDING_MAGIC 0x46564742u /* "FVGB" */

struct INFO {
  unsigned magic;
  RUN *run;
  fvg_ctx *ctx;
  fvh_flat *flat;
  int compat;
  int device;
  int status; /* of the last RHS / RK / faceinterp / gradcalc call */
};

typedef struct INFO fv_binding;

static fv_binding *find(RUN *run) {
  if (!run || !run->info) return NULL;
  fv_binding *b = run->info;
  return (b->magic == FV_BINDING_MAGIC && b->run == run) ? b : NULL;
}

static int leaves_io(RUN *run, int which, double *buf, int to_leaves) {
  const int neq = run->con->neq;
  int N = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt) ++N;
  int i = 0;
  for (BRANCH *b = run->topo->locl; b; b = b->lnxt, ++i) {
    LEAF *el = b->el;
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

static size_t field_len(RUN *run, int which, int N) {
  const int neq = run->con->neq;
  switch (which) {
    case FVG_FIELD_SF: return (size_t)6 * neq * N;
    case FVG_FIELD_SG: return (size_t)3 * (neq + 1) * N;
    default: return (size_t)neq * N;
  }
}

static int pull(fv_binding *b, int which) {
  int N = fvg_ncell(b->ctx);
  double *buf = (double *)malloc(field_len(b->run, which, N) * sizeof(double));
  if (!buf) return FVG_ERR_NOMEM;
  int rc = fvg_get_field(b->ctx, which, buf);
  if (!rc) rc = leaves_io(b->run, which, buf, 1);
  free(buf);
  return rc;
}

static int push(fv_binding *b, int which) {
  int N = fvg_ncell(b->ctx);
  double *buf = (double *)malloc(field_len(b->run, which, N) * sizeof(double));
  if (!buf) return FVG_ERR_NOMEM;
  int rc = leaves_io(b->run, which, buf, 0);
  if (!rc) rc = fvg_set_field(b->ctx, which, buf);
  free(buf);
  return rc;
}

int fv_gpu_attach(RUN *run, int device, const fvg_params *p, int compat) {
  if (!run || !p || !run->con || !run->topo) return FVG_ERR_ARG;
  /* con->iprimtv selects the MUSCL variables as in the reference (deltavalues.c:16).
   * With 1 the primitive path runs in its defined form: the reference's cons2primtv
   * divides by a density that facevalues_MUSCL never sets (cons2primtv.c:10-13), so its
   * own output is not a function of its inputs */
  if (run->con->iprimtv != 0 && run->con->iprimtv != 1) return FVG_ERR_UNSUPPORTED;
  fvg_params q = *p;
  q.iprimtv = run->con->iprimtv;
  fv_binding *b = (fv_binding *)calloc(1, sizeof(*b));
  if (!b) return FVG_ERR_NOMEM;
  int rc = fvg_create(device, &q, &b->ctx);
  if (rc) { free(b); return rc; }
  b->magic = FV_BINDING_MAGIC;
  b->run = run;
  b->compat = compat;
  b->device = device;
  run->info = b;
  rc = fv_gpu_remesh(run);
  if (rc) {
    fvg_destroy(b->ctx);
    b->magic = 0;
    free(b);
    run->info = NULL;
  }
  return rc;
}

int fv_gpu_detach(RUN *run) {
  fv_binding *b = find(run);
  if (!b) return FVG_ERR_STATE;
  fvg_destroy(b->ctx);
  fvh_flat_destroy(b->flat);
  b->magic = 0;
  free(b);
  run->info = NULL;
  return FVG_OK;
}

int fv_gpu_remesh(RUN *run) {
  fv_binding *b = find(run);
  if (!b) return FVG_ERR_STATE;
  fvh_flat *fl = NULL;
  int rc = fvh_flatten_run(run, &fl);
  if (rc) return rc;
  rc = fvg_set_mesh(b->ctx, fvh_flat_mesh(fl));
  if (rc) { fvh_flat_destroy(fl); return rc; }
  fvh_flat_destroy(b->flat);
  b->flat = fl;
  rc = push(b, FVG_FIELD_S);
  if (!rc) rc = push(b, FVG_FIELD_SN);
  return rc;
}

int fv_gpu_upload(RUN *run) {
  fv_binding *b = find(run);
  return b ? push(b, FVG_FIELD_S) : FVG_ERR_STATE;
}

int fv_gpu_download(RUN *run) {
  fv_binding *b = find(run);
  if (!b) return FVG_ERR_STATE;
  int rc = pull(b, FVG_FIELD_S);
  if (!rc) rc = pull(b, FVG_FIELD_SN);
  if (!rc) rc = pull(b, FVG_FIELD_RHS);
  return rc;
}

fvg_ctx *fv_gpu_context(RUN *run) {
  fv_binding *b = find(run);
  return b ? b->ctx : NULL;
}

int fv_gpu_last_status(RUN *run) {
  fv_binding *b = find(run);
  return b ? b->status : FVG_ERR_STATE;
}

void RHS(struct RUN *run) {
  fv_binding *b = find(run);
  if (!b) return;
  int rc = fvg_rhs(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_RHS);
  b->status = rc;
}

void RK(int irk, RUN *run) {
  fv_binding *b = find(run);
  if (!b) return;
  int rc = fvg_rk(b->ctx, irk, run->con->dt);
  if (!rc && b->compat) rc = pull(b, irk == 0 ? FVG_FIELD_SN : FVG_FIELD_S);
  b->status = rc;
}

void faceinterp(RUN *run) {
  fv_binding *b = find(run);
  if (!b) return;
  int rc = fvg_faceinterp(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_SF);
  b->status = rc;
}

void gradcalc(struct RUN *run) {
  fv_binding *b = find(run);
  if (!b) return;
  int rc = fvg_gradcalc(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_SG);
  b->status = rc;
}
