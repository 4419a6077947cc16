/* Parameter defaults: the reference's hard-coded constants. */
#include <string.h>

#include "fvgpu.h"

int fvg_neq(int model) {
  /* inputset.c:130-222: models 1 and 3 carry 5 equations, model 2 four */
  switch (model) {
    case 1: return 5;
    case 2: return 4;
    case 3: return 5;
    default: return -1;
  }
}

void fvg_params_default(fvg_params *p, int model) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->model = model;
  p->ns = 0;
  p->ischeme = 2;
  p->limiter = FVG_MINMOD;
  p->precision = FVG_PLAIN;
  /* con->eos gates the viscous work (flux_viscous.c:227) and is never set; CON is
   * malloc'd at ~2 MB (createrunstruct.c:27), i.e. fresh zero pages, so it reads 0 */
  p->visc_work = 0;
  p->gamma = 1.4;
  p->pinf = 0.0;
  p->mu = 0.0;
  p->ymin = 1e-6;
  /* barotropic_eos.c:8-12 */
  p->baro_nexp = 7.15;
  p->baro_beta = 293526643.0;
  p->baro_cbar = 1450.0;
  p->baro_psat = 2339.0;
  p->baro_rsat = 998.2;
  /* barotropic_eos.c:40-46 */
  p->mix_RG = 286.0;
  p->mix_Tref = 293.15;
  p->mix_rsat = 998.1618;
  p->mix_psat = 2340.0;
  p->mix_cliq = 1482.5;
  p->mix_cmix = 1.0;
  p->vanalbada_eps = 1e-12;
}
