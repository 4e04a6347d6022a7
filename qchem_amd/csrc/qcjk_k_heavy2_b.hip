// qcjk_k_heavy2_b.hip -- explicit instantiations of the register-transform wavefront kernel launchers (both bra transforms, both LDS batch sizes) for bra class dp.
// One translation unit per group so the kernel templates compile in parallel.
#include "qcjk_kernels.h"

namespace qcjk {

#define QCJK_INST(a, b, c, d)                                                                         \
    template hipError_t launch_heavy2<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2f<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2s<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2fs<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t); \
    template hipError_t launch_heavy2k<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);  \
    template hipError_t launch_heavy2ks<a, b, c, d>(const BuildArgs&, const ClassPairArgs&, hipStream_t);
QCJK_INST(2, 1, 0, 0) QCJK_INST(2, 1, 1, 0) QCJK_INST(2, 1, 1, 1) QCJK_INST(2, 1, 2, 0) QCJK_INST(2, 1, 2, 1)

}  // namespace qcjk
