// qcjk_k_heavy2r_b.hip -- explicit instantiations of the lane-R wavefront kernel launchers (both R batch
// sizes) for the classes with L = 5.
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d)                                                                          \
    template hipError_t launch_heavy2r<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2rw<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(2, 1, 1, 1) QCJK_INST(2, 1, 2, 0) QCJK_INST(2, 2, 1, 0)

}  // namespace qcjk
