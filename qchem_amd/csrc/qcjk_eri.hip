// qcjk_eri.hip -- dispatch table of the J/K class kernels: for each of the 21 bra-class >=
// ket-class combinations, the launchers of every kernel variant compiled for it.  The kernels
// themselves are in qcjk_kernels.h, instantiated by the qcjk_k_*.hip translation units.
#include "qcjk_launch.h"

namespace qcjk {

template <int LA, int LB, int LC, int LD>
constexpr ClassKernelInfo make_info()
{
    static_assert(has_light(LA, LB, LC, LD) == light_listed(LA, LB, LC, LD), "lane-kernel class list out of date");
    static_assert(has_lightc(LA, LB, LC, LD) == lightc_listed(LA, LB, LC, LD), "chunked lane-kernel class list out of date");
    ClassLauncher light = nullptr, lightc = nullptr;
    if constexpr (has_light(LA, LB, LC, LD)) light = &launch_light<LA, LB, LC, LD>;
    if constexpr (has_lightc(LA, LB, LC, LD)) lightc = &launch_lightc<LA, LB, LC, LD>;
    return {LA, LB, LC, LD, default_light(LA, LB, LC, LD) ? 64 : 1, light,
            &launch_heavy<LA, LB, LC, LD>, &launch_heavy2<LA, LB, LC, LD>, lightc, &launch_heavy2f<LA, LB, LC, LD>,
            &launch_heavy2s<LA, LB, LC, LD>, &launch_heavy2fs<LA, LB, LC, LD>, &launch_heavy2k<LA, LB, LC, LD>,
            &launch_heavy2ks<LA, LB, LC, LD>, &launch_heavy2r<LA, LB, LC, LD>, &launch_heavy2rw<LA, LB, LC, LD>};
}
#define QCJK_CLASS(LA, LB, LC, LD) make_info<LA, LB, LC, LD>()

// [bra class][ket class], bra >= ket; pair classes ss ps pp ds dp dd.
static const ClassKernelInfo g_table[kNumPairClasses][kNumPairClasses] = {
    {QCJK_CLASS(0, 0, 0, 0)},
    {QCJK_CLASS(1, 0, 0, 0), QCJK_CLASS(1, 0, 1, 0)},
    {QCJK_CLASS(1, 1, 0, 0), QCJK_CLASS(1, 1, 1, 0), QCJK_CLASS(1, 1, 1, 1)},
    {QCJK_CLASS(2, 0, 0, 0), QCJK_CLASS(2, 0, 1, 0), QCJK_CLASS(2, 0, 1, 1), QCJK_CLASS(2, 0, 2, 0)},
    {QCJK_CLASS(2, 1, 0, 0), QCJK_CLASS(2, 1, 1, 0), QCJK_CLASS(2, 1, 1, 1), QCJK_CLASS(2, 1, 2, 0),
     QCJK_CLASS(2, 1, 2, 1)},
    {QCJK_CLASS(2, 2, 0, 0), QCJK_CLASS(2, 2, 1, 0), QCJK_CLASS(2, 2, 1, 1), QCJK_CLASS(2, 2, 2, 0),
     QCJK_CLASS(2, 2, 2, 1), QCJK_CLASS(2, 2, 2, 2)},
};

const ClassKernelInfo* class_kernel(int bra_class, int ket_class)
{
    if (bra_class < 0 || bra_class >= kNumPairClasses || ket_class < 0 || ket_class > bra_class) return nullptr;
    return &g_table[bra_class][ket_class];
}

}  // namespace qcjk
