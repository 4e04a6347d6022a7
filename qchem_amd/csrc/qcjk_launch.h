// qcjk_launch.h -- per-class kernel launchers: declarations and the class lists each kernel
// translation unit instantiates.  The definitions live in qcjk_kernels.h; the qcjk_k_*.hip
// translation units instantiate them explicitly, and the dispatch table in qcjk_eri.hip links
// against those instantiations.
#pragma once

#include "qcjk_internal.h"

namespace qcjk {

constexpr int ncart_l(int l) { return (l + 1) * (l + 2) / 2; }

// Classes with a lane-per-quartet variant compiled (total angular momentum <= 4, block <= 81).
constexpr bool has_light(int la, int lb, int lc, int ld)
{
    return ncart_l(la) * ncart_l(lb) * ncart_l(lc) * ncart_l(ld) <= 81 && la + lb + lc + ld <= 4;
}
// Default choice before tuning (measured on the random-cluster workload): lane per quartet for
// L <= 4 except (dd|ss), the lane-R wavefront kernel with narrow batches (variant 10) otherwise.
constexpr bool default_light(int la, int lb, int lc, int ld)
{
    return has_light(la, lb, lc, ld) && !(la == 2 && lb == 2);
}
// Classes with the chunked lane-per-quartet variant: the lane-kernel classes with a p or d ket, and
// (dp|ds), whose per-chunk state still fits a lane without scratch ((dp|pp), (dd|ps) and (dd|pp)
// spill 1-5 KB per lane to scratch in FP64 and are not compiled).
constexpr bool has_lightc(int la, int lb, int lc, int ld)
{
    return (has_light(la, lb, lc, ld) && lc > 0) || (la == 2 && lb == 1 && lc == 2 && ld == 0);
}
// Bra Hermite tables go to LDS once they would cost more registers than the rest of the lane's state.
constexpr bool bra_in_lds(int la, int lb) { return la + lb >= 3; }

// The 21 bra-class >= ket-class combinations for l <= 2, and the subsets with lane kernels.
#define QCJK_ALL_CLASSES(X)                                                                              \
    X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0)        \
    X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0)        \
    X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)
#define QCJK_LIGHT_CLASSES(X)                                                                            \
    X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0)        \
    X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 2, 0, 0)
#define QCJK_LIGHTC_CLASSES(X)                                                                           \
    X(1, 0, 1, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 1, 0)        \
    X(2, 1, 2, 0)

#define QCJK_CLASS_MATCH(a, b, c, d) || (la == a && lb == b && lc == c && ld == d)
constexpr bool light_listed(int la, int lb, int lc, int ld) { return false QCJK_LIGHT_CLASSES(QCJK_CLASS_MATCH); }
constexpr bool lightc_listed(int la, int lb, int lc, int ld) { return false QCJK_LIGHTC_CLASSES(QCJK_CLASS_MATCH); }

template <int LA, int LB, int LC, int LD>
hipError_t launch_light(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_lightc(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2f(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2s(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2fs(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2k(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2ks(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2r(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2rw(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s);

}  // namespace qcjk
