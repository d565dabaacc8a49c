// pair_eam_alloy_hip.h -- "pair_style eam/alloy/hip": PairEAMAlloy whose
// compute() runs on the GPU through the C ABI (include/cac_hip.h).  The host
// tables stay for init_one; the device builds its own lists, so the neighbor
// request is an occasional one the reference never fills.
#ifdef PAIR_CLASS

PairStyle(eam/alloy/hip, PairEAMAlloyHIP)

#else

#ifndef CAC_PAIR_EAM_ALLOY_HIP_H
#define CAC_PAIR_EAM_ALLOY_HIP_H

#include "cac_hip.h"
#include "pair_eam_alloy.h"

namespace CAC_NS {

class PairEAMAlloyHIP : public PairEAMAlloy {
 public:
  PairEAMAlloyHIP(class CAC *);
  virtual ~PairEAMAlloyHIP() {}
  void coeff(int, char **);   // pair_eam_alloy.cpp:41-98 on the host, the setfl file on the device
  void init_style();          // pair_eam.cpp:310-320: upload and list settings instead of a list request
  void compute(int, int);     // pair_eam.cpp:618-1032 + reverse comm + Element::compute_nodef

 protected:
  cac_hip_ctx *ctx;
};

}  // namespace CAC_NS

#endif
#endif
