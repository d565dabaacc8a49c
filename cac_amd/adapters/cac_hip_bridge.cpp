// cac_hip_bridge.cpp -- see cac_hip_bridge.h
#include "cac_hip_bridge.h"

#include <cstring>

#include "atom.h"
#include "cac.h"
#include "domain.h"
#include "element.h"
#include "element_vec.h"
#include "error.h"
#include "update.h"

namespace CAC_NS {

cac_hip_ctx *hip_context(CAC *cac) {
  static cac_hip_ctx *ctx = nullptr;
  if (!ctx) {
    const int units = std::strcmp(cac->update->unit_style, "lj") == 0 ? CAC_HIP_UNITS_LJ : CAC_HIP_UNITS_METAL;
    // one rank per GPU: the rank's visible device is device 0
    if (cac_hip_create(&ctx, 0, units) != CAC_HIP_OK) cac->error->all(FLERR, "cac_hip_create failed");
  }
  return ctx;
}

void hip_check(Error *error, cac_hip_ctx *ctx, int rc) {
  if (rc != CAC_HIP_OK) error->all(FLERR, cac_hip_last_error(ctx));  // error.cpp:74-109
}

void hip_upload_model(CAC *cac, cac_hip_ctx *ctx) {
  Atom *atom = cac->atom;
  Element *element = cac->element;
  Error *error = cac->error;
  hip_check(error, ctx, cac_hip_set_box(ctx, cac->domain->boxlo, cac->domain->boxhi));
  // atom->mass is indexed by type, 1..ntypes
  hip_check(error, ctx, cac_hip_set_masses(ctx, atom->ntypes, atom->mass + 1));
  if (element->nlocal) {
    int **nc = element->evec->ncells, **ng = element->evec->ngcells;
    hip_check(error, ctx, cac_hip_add_etype(ctx, 1, nc[1][0], nc[1][1], nc[1][2], ng[1][0], ng[1][1], ng[1][2]));
  }
  // x [nlocal][3], v [nlocal][3]; nodex, nodev [nlocal][1][8][3] (one basis atom)
  hip_check(error, ctx, cac_hip_upload_atoms(ctx, atom->nlocal, atom->nlocal ? atom->x[0] : nullptr,
                                             atom->nlocal ? atom->v[0] : nullptr, atom->type));
  if (element->nlocal)
    hip_check(error, ctx, cac_hip_upload_elements(ctx, element->nlocal, element->etype, element->ctype[0],
                                                  element->nodex[0][0][0], element->nodev[0][0][0]));
}

void hip_download_state(CAC *cac, cac_hip_ctx *ctx) {
  Atom *atom = cac->atom;
  Element *element = cac->element;
  if (atom->nlocal)
    hip_check(cac->error, ctx, cac_hip_download_atoms(ctx, atom->x[0], atom->v[0], nullptr));
  if (element->nlocal)
    hip_check(cac->error, ctx, cac_hip_download_elements(ctx, element->nodex[0][0][0], element->nodev[0][0][0],
                                                         nullptr, nullptr));
}

}  // namespace CAC_NS
