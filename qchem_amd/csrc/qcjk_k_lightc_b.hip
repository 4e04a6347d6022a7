// qcjk_k_lightc_b.hip -- explicit instantiations of launch_lightc (lane-per-quartet kernel chunked over the ket
// first component) for (dp|ds).
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d) \
    template hipError_t launch_lightc<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(2, 1, 2, 0)

}  // namespace qcjk
