// cac_hip_bridge.h -- reference-side glue shared by the HIP adapters in this
// directory.  These files belong in the reference's src/ tree (they include
// its headers) and link against libcachip.so; this repository compiles them
// against the reference's headers only as a check (tests/test_adapters.py).
//
// One cac_hip_ctx per process (one rank per GPU), and the upload of the
// model from the arrays read_data fills (read_data.cpp:338-375): the box,
// the masses, the element type, atoms and elements.
#ifndef CAC_HIP_BRIDGE_H
#define CAC_HIP_BRIDGE_H

#include "cac_hip.h"

namespace CAC_NS {

class CAC;
class Error;

// the process's context, created on first use with the run's unit style
cac_hip_ctx *hip_context(CAC *cac);
// a non-zero return code is the reference's error->all with the library's message
void hip_check(Error *error, cac_hip_ctx *ctx, int rc);
// box, masses, element type, atoms and elements host -> device
void hip_upload_model(CAC *cac, cac_hip_ctx *ctx);
// atom x, v and element nodex, nodev device -> host
void hip_download_state(CAC *cac, cac_hip_ctx *ctx);

}  // namespace CAC_NS

#endif
