// write_data_hip.cpp -- see write_data_hip.h
#include "write_data_hip.h"

#include "cac_hip_bridge.h"
#include "error.h"

using namespace CAC_NS;

WriteDataHIP::WriteDataHIP(CAC *cac) : WriteData(cac) {}

void WriteDataHIP::command(int narg, char **arg) {
  cac_hip_ctx *ctx = hip_context(cac);
  hip_check(error, ctx, cac_hip_pbc(ctx));
  hip_download_state(cac, ctx);
  WriteData::command(narg, arg);  // its own Domain::pbc finds every point inside the box
}
