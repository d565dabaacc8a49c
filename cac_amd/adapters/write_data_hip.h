// write_data_hip.h -- the write_data command over a model that lives on the
// GPU: Domain::pbc on the device (as WriteData::command runs it on the host,
// write_data.cpp:120-131), the state copied back, then the reference's own
// writer.  The device lists are stale afterwards; the next run's setup
// rebuilds them, as the reference's does.
#ifndef CAC_WRITE_DATA_HIP_H
#define CAC_WRITE_DATA_HIP_H

#include "write_data.h"

namespace CAC_NS {

class WriteDataHIP : public WriteData {
 public:
  WriteDataHIP(class CAC *);
  void command(int, char **);
};

}  // namespace CAC_NS

#endif
