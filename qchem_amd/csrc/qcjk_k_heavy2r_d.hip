// qcjk_k_heavy2r_d.hip -- explicit instantiations of the lane-R wavefront kernel launchers (both R batch
// sizes) for the classes with L = 7 and 8 (R tables built in place in LDS).
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d)                                                                          \
    template hipError_t launch_heavy2r<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2rw<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(2, 2, 2, 1) QCJK_INST(2, 2, 2, 2)

}  // namespace qcjk
