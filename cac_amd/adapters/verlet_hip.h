// verlet_hip.h -- "run_style verlet/hip": Verlet::setup and Verlet::run with
// the whole step loop on the GPU (cac_hip_setup, cac_hip_run): FixNVE, the
// rebuild trigger, pbc, lists, forces, nodal assembly and thermo, with the
// arrays copied back only when the run ends.
#ifdef INTEGRATE_CLASS

IntegrateStyle(verlet/hip, VerletHIP)

#else

#ifndef CAC_VERLET_HIP_H
#define CAC_VERLET_HIP_H

#include "verlet.h"

namespace CAC_NS {

class VerletHIP : public Verlet {
 public:
  VerletHIP(class CAC *, int, char **);
  virtual ~VerletHIP() {}
  void setup(int, int flag = 1);  // verlet.cpp:65-186
  void run(int, int);             // verlet.cpp:218-360
};

}  // namespace CAC_NS

#endif
#endif
