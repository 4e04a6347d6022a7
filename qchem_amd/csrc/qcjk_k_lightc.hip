// qcjk_k_lightc.hip -- explicit instantiations of launch_lightc (lane-per-quartet kernel chunked over the ket
// first component) for the lane-kernel classes with a p or d ket.
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d) \
    template hipError_t launch_lightc<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(1, 0, 1, 0) QCJK_INST(1, 1, 1, 0) QCJK_INST(1, 1, 1, 1) QCJK_INST(2, 0, 1, 0) QCJK_INST(2, 0, 1, 1)
QCJK_INST(2, 0, 2, 0) QCJK_INST(2, 1, 1, 0)

}  // namespace qcjk
