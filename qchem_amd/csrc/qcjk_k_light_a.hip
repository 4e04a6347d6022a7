// qcjk_k_light_a.hip -- explicit instantiations of launch_light (lane-per-quartet kernel) for the classes with L <= 3.
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d) \
    template hipError_t launch_light<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(0, 0, 0, 0) QCJK_INST(1, 0, 0, 0) QCJK_INST(1, 0, 1, 0) QCJK_INST(1, 1, 0, 0) QCJK_INST(1, 1, 1, 0)
QCJK_INST(2, 0, 0, 0) QCJK_INST(2, 0, 1, 0) QCJK_INST(2, 1, 0, 0)

}  // namespace qcjk
