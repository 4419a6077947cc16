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
 * The reference reports errors by printf + exit(0) (residual_faces_AMR.c:186-194);
 * these wrappers keep the void signature and record the status instead
 * (fv_gpu_last_status).
 */
#include <stdlib.h>
#include <string.h>

#include "fvgpu.h"

#define FV_MAX_BIND 64

typedef struct {
  RUN *run;
  fvg_ctx *ctx;
  fvh_flat *flat;
  int compat;
  int device;
} fv_binding;

static fv_binding g_bind[FV_MAX_BIND];
static int g_status = FVG_OK;

static fv_binding *find(RUN *run) {
  for (int i = 0; i < FV_MAX_BIND; ++i)
    if (g_bind[i].run == run && run) return &g_bind[i];
  return NULL;
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
  if (!run || !p) return FVG_ERR_ARG;
  if (find(run)) return FVG_ERR_STATE;
  /* MUSCL on primitive variables: the reference's cons2primtv divides by a density
   * that facevalues_MUSCL never sets (cons2primtv.c:10-13), so there is no defined
   * result to reproduce */
  if (run->con && run->con->iprimtv != 0) return FVG_ERR_UNSUPPORTED;
  fv_binding *b = NULL;
  for (int i = 0; i < FV_MAX_BIND; ++i)
    if (!g_bind[i].run) { b = &g_bind[i]; break; }
  if (!b) return FVG_ERR_NOMEM;
  memset(b, 0, sizeof(*b));
  int rc = fvg_create(device, p, &b->ctx);
  if (rc) return rc;
  b->run = run;
  b->compat = compat;
  b->device = device;
  rc = fv_gpu_remesh(run);
  if (rc) {
    fvg_destroy(b->ctx);
    memset(b, 0, sizeof(*b));
  }
  return rc;
}

int fv_gpu_detach(RUN *run) {
  fv_binding *b = find(run);
  if (!b) return FVG_ERR_STATE;
  fvg_destroy(b->ctx);
  fvh_flat_destroy(b->flat);
  memset(b, 0, sizeof(*b));
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

int fv_gpu_last_status(void) { return g_status; }

void RHS(struct RUN *run) {
  fv_binding *b = find(run);
  if (!b) { g_status = FVG_ERR_STATE; return; }
  int rc = fvg_rhs(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_RHS);
  g_status = rc;
}

void RK(int irk, RUN *run) {
  fv_binding *b = find(run);
  if (!b) { g_status = FVG_ERR_STATE; return; }
  int rc = fvg_rk(b->ctx, irk, run->con->dt);
  if (!rc && b->compat) rc = pull(b, irk == 0 ? FVG_FIELD_SN : FVG_FIELD_S);
  g_status = rc;
}

void faceinterp(RUN *run) {
  fv_binding *b = find(run);
  if (!b) { g_status = FVG_ERR_STATE; return; }
  int rc = fvg_faceinterp(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_SF);
  g_status = rc;
}

void gradcalc(struct RUN *run) {
  fv_binding *b = find(run);
  if (!b) { g_status = FVG_ERR_STATE; return; }
  int rc = fvg_gradcalc(b->ctx);
  if (!rc && b->compat) rc = pull(b, FVG_FIELD_SG);
  g_status = rc;
}
