// pair_eam_alloy_hip.cpp -- see pair_eam_alloy_hip.h
#include "pair_eam_alloy_hip.h"

#include "atom.h"
#include "cac_hip_bridge.h"
#include "element.h"
#include "error.h"
#include "neigh_request.h"
#include "neighbor.h"

using namespace CAC_NS;

// PairEAMAlloy derives virtually from PairEAM, so the most derived class
// constructs PairEAM
PairEAMAlloyHIP::PairEAMAlloyHIP(CAC *cac) : PairEAM(cac), PairEAMAlloy(cac), ctx(nullptr) {}

void PairEAMAlloyHIP::coeff(int narg, char **arg) {
  PairEAMAlloy::coeff(narg, arg);  // host tables, setflag, element map
  ctx = hip_context(cac);
  hip_check(error, ctx, cac_hip_pair_eam_setfl(ctx, arg[2]));
}

void PairEAMAlloyHIP::init_style() {
  int irequest = neighbor->request(this, instance_me);
  neighbor->requests[irequest]->occasional = 1;
  hip_upload_model(cac, ctx);
  hip_check(error, ctx, cac_hip_neighbor_settings(ctx, neighbor->skin, neighbor->every, neighbor->delay,
                                                  neighbor->dist_check, neighbor->oneatom));
  hip_check(error, ctx, cac_hip_neigh_build(ctx));
}

void PairEAMAlloyHIP::compute(int eflag, int vflag) {
  ev_init(eflag, vflag);
  // atom->x [nlocal][3] and element->nodex [nlocal][1][8][3] in, as the reference stores them
  hip_check(error, ctx, cac_hip_sync_positions(ctx, atom->nlocal ? atom->x[0] : nullptr,
                                               element->nlocal ? element->nodex[0][0][0] : nullptr));
  int rebuild = 0;
  hip_check(error, ctx, cac_hip_neigh_decide(ctx, &rebuild));
  if (rebuild) hip_check(error, ctx, cac_hip_neigh_build(ctx));
  hip_check(error, ctx, cac_hip_compute(ctx, eflag, vflag));
  hip_check(error, ctx, cac_hip_get_energy(ctx, &eng_vdwl, virial));
  // ev_setup allocated eatom/enode, vatom/vnode for ENERGY_ATOM / VIRIAL_ATOM
  // (or VIRIAL_CENTROID, pair.cpp:266-345); the device formed them
  if (eflag_atom || vflag_atom)
    hip_check(error, ctx, cac_hip_get_peratom(ctx, eflag_atom && atom->nlocal ? eatom : nullptr,
                                              eflag_atom && element->nlocal ? enode[0][0] : nullptr,
                                              vflag_atom && atom->nlocal ? vatom[0] : nullptr,
                                              vflag_atom && element->nlocal ? vnode[0][0][0] : nullptr));
  // atom->f, element->nodef (the assembly Element::compute_nodef forms,
  // element.cpp:2178-2246) and element->gaussf out
  hip_check(error, ctx, cac_hip_download_forces(ctx, atom->nlocal ? atom->f[0] : nullptr,
                                                element->nlocal ? element->nodef[0][0][0] : nullptr,
                                                element->nlocal ? element->gaussf[0][0][0] : nullptr));
}
