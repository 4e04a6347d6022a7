// qcjk_k_heavy.hip -- explicit instantiations of launch_heavy (LDS-staged wavefront kernel) for all 21 classes.
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d) \
    template hipError_t launch_heavy<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_ALL_CLASSES(QCJK_INST)

}  // namespace qcjk
