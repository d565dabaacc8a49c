// verlet_hip.cpp -- see verlet_hip.h
#include "verlet_hip.h"

#include <vector>

#include "cac.h"
#include "cac_hip_bridge.h"
#include "error.h"
#include "output.h"
#include "update.h"

using namespace CAC_NS;

VerletHIP::VerletHIP(CAC *cac, int narg, char **arg) : Verlet(cac, narg, arg) {}

void VerletHIP::setup(int /*forceflag*/, int /*flag*/) {
  cac_hip_ctx *ctx = hip_context(cac);
  hip_upload_model(cac, ctx);
  // neighbor build + the first force with energy and virial
  hip_check(error, ctx, cac_hip_setup(ctx, 1, 1));
}

void VerletHIP::run(int n, int /*forceflag*/) {
  cac_hip_ctx *ctx = hip_context(cac);
  const int every = output->thermo_every;
  // rows of (step, pe, ke, etotal, press): the first, every `every` steps, the last
  std::vector<double> th(5 * (size_t)(n / (every > 0 ? every : n > 0 ? n : 1) + 3));
  hip_check(error, ctx, cac_hip_run(ctx, n, update->dt, every, th.data()));
  hip_download_state(cac, ctx);
  update->ntimestep += n;
}
