// qcjk_kernels.h -- J/K class kernels for gfx950 and their launchers (templates).  Included only by
// the kernel translation units (qcjk_k_*.hip), each of which instantiates one kernel shape for its
// classes, so the heavy templates compile in parallel; the dispatch table (qcjk_eri.hip) sees
// only the declarations in qcjk_launch.h.
//
// Replaces, for the molecular Gaussian path, the reference's per-element ERI loop and ERI4
// contraction (Molecule/IrrepBasisSet.C:138-183, Internal/Imp/ERI4.C:10-42): every unique shell
// quartet (8-fold symmetry) is evaluated once and digested straight into J and K, with FP64
// hardware atomics into the unsymmetrised accumulators Jt, Kt:
//   s = (A==B ? 1/2 : 1)(C==D ? 1/2 : 1)(AB==CD ? 1/2 : 1)
//   Jt_ab += 2s sum_cd I D_cd   Jt_cd += 2s sum_ab I D_ab
//   Kt_ac += s sum_bd I D_bd    Kt_bc += s sum_ad I D_ad   Kt_ad += s sum_bc I D_bc   Kt_bd += s sum_ac I D_ac
// and J = Jt + Jt^T, K = Kt + Kt^T afterwards (qcjk_api.hip).
//
// Kernel shapes (the variant numbers of qcjk_set_variant):
//   0 k_jk_light   lane per contracted quartet.  A wavefront owns one bra pair and 64 consecutive
//                  ket pairs, so the J_ab contribution is reduced across the wavefront before one
//                  atomic per component; everything else is register resident.
//   1 k_jk_heavy   wavefront per quartet, LDS-staged transforms (mnd_heavy.h).
//   2 k_jk_heavy2  wavefront per quartet, register-resident transforms, batched R (mnd_heavy2.h).
//   3 k_jk_lightc  lane per quartet, chunked over the ket's first component.
//   4 k_jk_heavy2  with the sum-factorised bra transform (FACT = true).
//   5, 6           variants 2 and 4 with small LDS batches (FORM 1): more wavefronts per SIMD.
//   7, 8           k_jk_heavy2 with both transforms sum-factorised (FACT = 3), big / small batches.
//   9, 10          k_jk_heavy2 in lane-R form (FORM 2, 3): ~16 KB R batches / ~8 KB R batches and at
//                  most 256 registers.
// With QCJK_MIXED, variants 0, 2 and 3 evaluate small quartets' primitive integrals in FP32
// (Boys at T >= 40 stays FP64); the contracted block is digested into J/K in FP64.  The FP32 work
// runs in its own instantiation of the kernel (FP32 = true), launched after the FP64 one over the
// same work list: each takes only the units of its precision, so each keeps its own register
// allocation and neither carries the other's code.
#pragma once

#include <algorithm>

#include "qcjk_launch.h"
#include "mnd_quartet.h"
#include "mnd_heavy2.h"

namespace qcjk {

__device__ __forceinline__ int find_unit_owner(const long long* __restrict__ start, int n, long long u)
{
    int lo = 0, hi = n;  // start[lo] <= u < start[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Wavefront kernels (one quartet per unit): each block takes a contiguous run of units, so the
// owning bra pair is found by one binary search and then followed forward (a unit's bra is the
// previous unit's or a later one), instead of one dependent search per quartet.  Within a class
// every quartet has the same shape, so equal runs balance.
struct UnitRange {
    long long u, end, next;
    int P;
    __device__ __forceinline__ explicit UnitRange(const ClassPairArgs& C)
    {
        const long long per = (C.nunits + gridDim.x - 1) / gridDim.x;
        u = (long long)blockIdx.x * per;
        end = u + per < C.nunits ? u + per : C.nunits;
        P = 0;
        next = 0;
        if (u < end) {
            P = find_unit_owner(C.unit_start, C.nbra, u);
            next = C.unit_start[P + 1];
        }
    }
    // owner of unit v (v increases by one per call, v < nunits = unit_start[nbra])
    __device__ __forceinline__ int owner(const ClassPairArgs& C, long long v)
    {
        while (v >= next) next = C.unit_start[++P + 1];
        return P;
    }
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Mixed precision, lane-per-quartet kernels: a wavefront's unit belongs to the FP32 kernel when
// every active lane's quartet has Q_ab Q_cd max|D'| below mixed_tol (one ballot, so the choice is
// uniform over the wavefront).  Ket pairs are sorted by descending bound, so the choice is almost
// always the per-lane one.
__device__ __forceinline__ bool wave_fp32(const BuildArgs& A, bool active, double4 pg, double4 qg)
{
    if (!(A.flags & kFlagMixed)) return false;
    const bool big = active && !(pg.w * qg.w * (*A.dmax) < A.mixed_tol);
    return __ballot(big) == 0;
}

// Per-wavefront tallies of the quartets computed (all, FP32, primitive quartets), added to the
// class's counters once when the wavefront ends: atomics per quartet on three global words would
// queue in one L2 channel.  Primitive counts use lane 0's contraction lengths.
struct WaveTally {
    unsigned long long q = 0, q32 = 0, prim = 0;
    __device__ __forceinline__ void lanes(bool active, bool fp32, int4 pi, int4 qi)  // every lane calls
    {
        const unsigned long long cnt = (unsigned long long)__popcll(__ballot(active));
        q += cnt;
        if (fp32) q32 += cnt;
        prim += cnt * (unsigned long long)((pi.w & 0xffff) * (qi.w & 0xffff));
    }
    __device__ __forceinline__ void one(bool fp32, int4 pi, int4 qi)
    {
        q += 1;
        if (fp32) q32 += 1;
        prim += (unsigned long long)((pi.w & 0xffff) * (qi.w & 0xffff));
    }
    __device__ __forceinline__ void flush(const BuildArgs& A, int lane) const
    {
        if (A.counters && lane == 0 && q) {
            atomicAdd(&A.counters[0], q);
            if (q32) atomicAdd(&A.counters[1], q32);
            atomicAdd(&A.counters[2], prim);
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Lane-per-quartet kernel
// ---------------------------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, bool BRA_MEM, bool FP32>
__global__ __launch_bounds__(256) void k_jk_light(BuildArgs A, ClassPairArgs C)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    using RT = std::conditional_t<FP32, float, double>;
    __shared__ HermiteAxis<LA, LB, RT> sEab[BRA_MEM ? 4 : 1][3];
    const int lane = threadIdx.x & 63;
    WaveTally tally;
    for (long long unit = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); unit < C.nunits;
         unit += (long long)gridDim.x * 4) {  // uniform per wavefront
        const int P = find_unit_owner(C.unit_start, C.nbra, unit);
        const int q = (int)((unit - C.unit_start[P]) * 64) + lane;
        const int2 seg = C.kseg[P];
        bool active = q < seg.y;
        const int Q = seg.x + (active ? q : 0);
        const int2 mm = A.pmat[C.bra_off + P];
        const int n = mm.y;
        const double* __restrict__ D = A.D + mm.x;
        double* __restrict__ Jt = A.Jt + mm.x;
        double* __restrict__ Kt = A.Kt + mm.x;
        const int4 pi = A.pinfo[C.bra_off + P];
        const double4 pg = A.pgeo[C.bra_off + P];
        const int4 qi = A.pinfo[C.ket_off + Q];
        const double4 qg = A.pgeo[C.ket_off + Q];
        if (A.flags & kFlagScreen) active = active && (pg.w * qg.w * (*A.dmax) >= A.screen_tol);
        if (wave_fp32(A, active, pg, qg) != FP32) continue;  // the other precision's unit
        const int slot = BRA_MEM ? (threadIdx.x >> 6) : 0;

        double I[NAB * NCD];
        static_for<0, NAB * NCD>([&](auto i) { I[decltype(i)::value] = 0.0; });
        if (active) {
            if constexpr (FP32) {
                float If[NAB * NCD];
                static_for<0, NAB * NCD>([&](auto i) { If[decltype(i)::value] = 0.0f; });
                light_compute<LA, LB, LC, LD, BRA_MEM, float>(A, pi, pg, qi, qg, If, sEab[slot]);
                static_for<0, NAB * NCD>([&](auto i) { I[decltype(i)::value] = (double)If[decltype(i)::value]; });
            } else {
                light_compute<LA, LB, LC, LD, BRA_MEM, double>(A, pi, pg, qi, qg, I, sEab[slot]);
            }
        }

        double s = ((pi.w >> 16) ? 0.5 : 1.0) * ((qi.w >> 16) ? 0.5 : 1.0) * ((C.same && Q == P) ? 0.5 : 1.0);
        if (!active) s = 0.0;
        const int oa = pi.x, ob = pi.y, oc = qi.x, od = qi.y;

        if (A.flags & kFlagJ) {
            double dcd[NCD];
            static_for<0, NCD>([&](auto i) {
                constexpr int c = decltype(i)::value / ND, d = decltype(i)::value % ND;
                dcd[decltype(i)::value] = active ? D[(oc + c) * n + od + d] : 0.0;
            });
            static_for<0, NAB>([&](auto ab) {
                constexpr int a = decltype(ab)::value / NB, b = decltype(ab)::value % NB;
                double v = 0.0;
                static_for<0, NCD>([&](auto cd) { v += I[decltype(ab)::value * NCD + decltype(cd)::value] * dcd[decltype(cd)::value]; });
                v = wave_sum(2.0 * s * v);
                if (lane == 0) unsafeAtomicAdd(&Jt[(oa + a) * n + ob + b], v);
            });
            double dab[NAB];
            static_for<0, NAB>([&](auto i) {
                constexpr int a = decltype(i)::value / NB, b = decltype(i)::value % NB;
                dab[decltype(i)::value] = D[(oa + a) * n + ob + b];
            });
            if (active) {
                static_for<0, NCD>([&](auto cd) {
                    constexpr int c = decltype(cd)::value / ND, d = decltype(cd)::value % ND;
                    double v = 0.0;
                    static_for<0, NAB>([&](auto ab) { v += I[decltype(ab)::value * NCD + decltype(cd)::value] * dab[decltype(ab)::value]; });
                    unsafeAtomicAdd(&Jt[(oc + c) * n + od + d], 2.0 * s * v);
                });
            }
        }
        if ((A.flags & kFlagK) && active) {
            // Kt_ac += s sum_bd I D_bd
            {
                double dd[NB * ND];
                static_for<0, NB * ND>([&](auto i) {
                    constexpr int b = decltype(i)::value / ND, d = decltype(i)::value % ND;
                    dd[decltype(i)::value] = D[(ob + b) * n + od + d];
                });
                static_for<0, NA * NC>([&](auto x) {
                    constexpr int a = decltype(x)::value / NC, c = decltype(x)::value % NC;
                    double v = 0.0;
                    static_for<0, NB * ND>([&](auto y) {
                        constexpr int b = decltype(y)::value / ND, d = decltype(y)::value % ND;
                        v += I[((a * NB + b) * NC + c) * ND + d] * dd[decltype(y)::value];
                    });
                    unsafeAtomicAdd(&Kt[(oa + a) * n + oc + c], s * v);
                });
            }
            // Kt_bc += s sum_ad I D_ad
            {
                double dd[NA * ND];
                static_for<0, NA * ND>([&](auto i) {
                    constexpr int a = decltype(i)::value / ND, d = decltype(i)::value % ND;
                    dd[decltype(i)::value] = D[(oa + a) * n + od + d];
                });
                static_for<0, NB * NC>([&](auto x) {
                    constexpr int b = decltype(x)::value / NC, c = decltype(x)::value % NC;
                    double v = 0.0;
                    static_for<0, NA * ND>([&](auto y) {
                        constexpr int a = decltype(y)::value / ND, d = decltype(y)::value % ND;
                        v += I[((a * NB + b) * NC + c) * ND + d] * dd[decltype(y)::value];
                    });
                    unsafeAtomicAdd(&Kt[(ob + b) * n + oc + c], s * v);
                });
            }
            // Kt_ad += s sum_bc I D_bc
            {
                double dd[NB * NC];
                static_for<0, NB * NC>([&](auto i) {
                    constexpr int b = decltype(i)::value / NC, c = decltype(i)::value % NC;
                    dd[decltype(i)::value] = D[(ob + b) * n + oc + c];
                });
                static_for<0, NA * ND>([&](auto x) {
                    constexpr int a = decltype(x)::value / ND, d = decltype(x)::value % ND;
                    double v = 0.0;
                    static_for<0, NB * NC>([&](auto y) {
                        constexpr int b = decltype(y)::value / NC, c = decltype(y)::value % NC;
                        v += I[((a * NB + b) * NC + c) * ND + d] * dd[decltype(y)::value];
                    });
                    unsafeAtomicAdd(&Kt[(oa + a) * n + od + d], s * v);
                });
            }
            // Kt_bd += s sum_ac I D_ac
            {
                double dd[NA * NC];
                static_for<0, NA * NC>([&](auto i) {
                    constexpr int a = decltype(i)::value / NC, c = decltype(i)::value % NC;
                    dd[decltype(i)::value] = D[(oa + a) * n + oc + c];
                });
                static_for<0, NB * ND>([&](auto x) {
                    constexpr int b = decltype(x)::value / ND, d = decltype(x)::value % ND;
                    double v = 0.0;
                    static_for<0, NA * NC>([&](auto y) {
                        constexpr int a = decltype(y)::value / NC, c = decltype(y)::value % NC;
                        v += I[((a * NB + b) * NC + c) * ND + d] * dd[decltype(y)::value];
                    });
                    unsafeAtomicAdd(&Kt[(ob + b) * n + od + d], s * v);
                });
            }
        }
        tally.lanes(active, FP32, pi, qi);
    }
    tally.flush(A, lane);
}

// ---------------------------------------------------------------------------------------------
// Lane-per-quartet kernel, chunked over the ket's first component: the primitive loop runs once
// per chunk (R recomputed), so only NAB*ND accumulators are live.  J_ab, K_ad, K_bd accumulate
// across chunks; J_cd, K_ac, K_bc of a chunk are final when it ends.
// ---------------------------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, bool BRA_MEM, bool FP32>
__global__ __launch_bounds__(256) void k_jk_lightc(BuildArgs A, ClassPairArgs C)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB;
    using RT = std::conditional_t<FP32, float, double>;
    __shared__ HermiteAxis<LA, LB, RT> sEab[BRA_MEM ? 4 : 1][3];
    const int lane = threadIdx.x & 63;
    WaveTally tally;
    for (long long unit = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); unit < C.nunits;
         unit += (long long)gridDim.x * 4) {  // uniform per wavefront
        const int P = find_unit_owner(C.unit_start, C.nbra, unit);
        const int q = (int)((unit - C.unit_start[P]) * 64) + lane;
        const int2 seg = C.kseg[P];
        bool active = q < seg.y;
        const int Q = seg.x + (active ? q : 0);
        const int2 mm = A.pmat[C.bra_off + P];
        const int n = mm.y;
        const double* __restrict__ D = A.D + mm.x;
        double* __restrict__ Jt = A.Jt + mm.x;
        double* __restrict__ Kt = A.Kt + mm.x;
        const int4 pi = A.pinfo[C.bra_off + P];
        const double4 pg = A.pgeo[C.bra_off + P];
        const int4 qi = A.pinfo[C.ket_off + Q];
        const double4 qg = A.pgeo[C.ket_off + Q];
        if (A.flags & kFlagScreen) active = active && (pg.w * qg.w * (*A.dmax) >= A.screen_tol);
        if (wave_fp32(A, active, pg, qg) != FP32) continue;  // the other precision's unit
        const int slot = BRA_MEM ? (threadIdx.x >> 6) : 0;
        double s = ((pi.w >> 16) ? 0.5 : 1.0) * ((qi.w >> 16) ? 0.5 : 1.0) * ((C.same && Q == P) ? 0.5 : 1.0);
        if (!active) s = 0.0;
        const int oa = pi.x, ob = pi.y, oc = qi.x, od = qi.y;
        const bool doJ = (A.flags & kFlagJ) != 0, doK = (A.flags & kFlagK) != 0;
        double jab[NAB], kad[NA * ND], kbd[NB * ND];
        static_for<0, NAB>([&](auto i) { jab[decltype(i)::value] = 0.0; });
        static_for<0, NA * ND>([&](auto i) { kad[decltype(i)::value] = 0.0; });
        static_for<0, NB * ND>([&](auto i) { kbd[decltype(i)::value] = 0.0; });
        static_for<0, NC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            double I[NAB * ND];
            static_for<0, NAB * ND>([&](auto i) { I[decltype(i)::value] = 0.0; });
            if (active) {
                if constexpr (FP32) {
                    float If[NAB * ND];
                    static_for<0, NAB * ND>([&](auto i) { If[decltype(i)::value] = 0.0f; });
                    light_compute_chunk<LA, LB, LC, LD, c, BRA_MEM, float>(A, pi, pg, qi, qg, If, sEab[slot]);
                    static_for<0, NAB * ND>([&](auto i) { I[decltype(i)::value] = (double)If[decltype(i)::value]; });
                } else {
                    light_compute_chunk<LA, LB, LC, LD, c, BRA_MEM, double>(A, pi, pg, qi, qg, I, sEab[slot]);
                }
            }
            if (active && doJ) {
                static_for<0, ND>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    const double dcd = D[(oc + c) * n + od + d];
                    double v = 0.0;
                    static_for<0, NAB>([&](auto ab) {
                        constexpr int a = decltype(ab)::value / NB, b = decltype(ab)::value % NB;
                        jab[decltype(ab)::value] += I[decltype(ab)::value * ND + d] * dcd;
                        v += I[decltype(ab)::value * ND + d] * D[(oa + a) * n + ob + b];
                    });
                    unsafeAtomicAdd(&Jt[(oc + c) * n + od + d], 2.0 * s * v);
                });
            }
            if (active && doK) {
                static_for<0, NA>([&](auto aa) {  // Kt_ac += s sum_bd I D_bd
                    constexpr int a = decltype(aa)::value;
                    double v = 0.0;
                    static_for<0, NB>([&](auto bb) {
                        constexpr int b = decltype(bb)::value;
                        static_for<0, ND>([&](auto dd) {
                            constexpr int d = decltype(dd)::value;
                            v += I[(a * NB + b) * ND + d] * D[(ob + b) * n + od + d];
                        });
                    });
                    unsafeAtomicAdd(&Kt[(oa + a) * n + oc + c], s * v);
                });
                static_for<0, NB>([&](auto bb) {  // Kt_bc += s sum_ad I D_ad
                    constexpr int b = decltype(bb)::value;
                    double v = 0.0;
                    static_for<0, NA>([&](auto aa) {
                        constexpr int a = decltype(aa)::value;
                        static_for<0, ND>([&](auto dd) {
                            constexpr int d = decltype(dd)::value;
                            v += I[(a * NB + b) * ND + d] * D[(oa + a) * n + od + d];
                        });
                    });
                    unsafeAtomicAdd(&Kt[(ob + b) * n + oc + c], s * v);
                });
                static_for<0, NA>([&](auto aa) {  // Kt_ad += s sum_bc I D_bc (this chunk's c)
                    constexpr int a = decltype(aa)::value;
                    static_for<0, ND>([&](auto dd) {
                        constexpr int d = decltype(dd)::value;
                        double v = 0.0;
                        static_for<0, NB>([&](auto bb) {
                            constexpr int b = decltype(bb)::value;
                            v += I[(a * NB + b) * ND + d] * D[(ob + b) * n + oc + c];
                        });
                        kad[a * ND + d] += v;
                    });
                });
                static_for<0, NB>([&](auto bb) {  // Kt_bd += s sum_ac I D_ac (this chunk's c)
                    constexpr int b = decltype(bb)::value;
                    static_for<0, ND>([&](auto dd) {
                        constexpr int d = decltype(dd)::value;
                        double v = 0.0;
                        static_for<0, NA>([&](auto aa) {
                            constexpr int a = decltype(aa)::value;
                            v += I[(a * NB + b) * ND + d] * D[(oa + a) * n + oc + c];
                        });
                        kbd[b * ND + d] += v;
                    });
                });
            }
        });
        if (doJ) {
            static_for<0, NAB>([&](auto ab) {
                constexpr int a = decltype(ab)::value / NB, b = decltype(ab)::value % NB;
                const double v = wave_sum(2.0 * s * jab[decltype(ab)::value]);
                if (lane == 0) unsafeAtomicAdd(&Jt[(oa + a) * n + ob + b], v);
            });
        }
        if (doK && active) {
            static_for<0, NA * ND>([&](auto x) {
                constexpr int a = decltype(x)::value / ND, d = decltype(x)::value % ND;
                unsafeAtomicAdd(&Kt[(oa + a) * n + od + d], s * kad[decltype(x)::value]);
            });
            static_for<0, NB * ND>([&](auto x) {
                constexpr int b = decltype(x)::value / ND, d = decltype(x)::value % ND;
                unsafeAtomicAdd(&Kt[(ob + b) * n + od + d], s * kbd[decltype(x)::value]);
            });
        }
        tally.lanes(active, FP32, pi, qi);
    }
    tally.flush(A, lane);
}

// ---------------------------------------------------------------------------------------------
// Wavefront-per-quartet kernel (grid-stride over quartets), LDS-staged transforms.  FP64 only:
// QCJK_MIXED leaves this variant's quartets in FP64.
// ---------------------------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_jk_heavy(BuildArgs A, ClassPairArgs C)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    using Dg = HeavyDigest<LA, LB, LC, LD>;
    __shared__ HeavySmem<LA, LB, LC, LD> sm;
    const int lane = threadIdx.x;
    HeavyLane<LA, LB, LC, LD> ln;
    ln.init(lane);
    const bool doJ = (A.flags & kFlagJ) != 0, doK = (A.flags & kFlagK) != 0;
    WaveTally tally;
    UnitRange ur(C);
    for (long long u = ur.u; u < ur.end; ++u) {
        const int P = ur.owner(C, u);
        const int Q = C.kseg[P].x + (int)(u - C.unit_start[P]);
        const int2 mm = A.pmat[C.bra_off + P];
        const int4 pi = A.pinfo[C.bra_off + P];
        const double4 pg = A.pgeo[C.bra_off + P];
        const int4 qi = A.pinfo[C.ket_off + Q];
        const double4 qg = A.pgeo[C.ket_off + Q];
        if ((A.flags & kFlagScreen) && !(pg.w * qg.w * (*A.dmax) >= A.screen_tol)) continue;
        heavy_compute<LA, LB, LC, LD>(A, sm, ln, lane, pi, pg, qi, qg);
        const double s = ((pi.w >> 16) ? 0.5 : 1.0) * ((qi.w >> 16) ? 0.5 : 1.0) * ((C.same && Q == P) ? 0.5 : 1.0);
        for (int t = lane; t < Dg::NTASK; t += 64) {
            int row, col;
            double val;
            if (Dg::task(t, sm.XI, A.D + mm.x, mm.y, pi.x, pi.y, qi.x, qi.y, doJ, doK, s, row, col, val)) {
                double* tgt = (t < Dg::T2 ? A.Jt : A.Kt) + mm.x;
                unsafeAtomicAdd(&tgt[row * mm.y + col], val);
            }
        }
        tally.one(false, pi, qi);
    }
    tally.flush(A, lane);
}

// ---------------------------------------------------------------------------------------------
// Wavefront-per-quartet kernel with register-resident transforms (mnd_heavy2.h)
// ---------------------------------------------------------------------------------------------
// FORM: 0 level-by-level R batches, 1 the same with small LDS batches, 2 lane-R form, 3 lane-R form
// with narrow R and X batches (~8 KB each) held to 256 registers (VGPRs + AGPRs): two wavefronts per
// SIMD.
template <int LA, int LB, int LC, int LD, bool FP32, int FACT, int FORM>
__global__ __launch_bounds__(64, (FORM == 3 ? 2 : 1)) void k_jk_heavy2(BuildArgs A, ClassPairArgs C)
{
    constexpr bool SMALL = FORM == 1, RL = FORM >= 2, NARROW = FORM == 3;
    using Dg = HeavyDigest<LA, LB, LC, LD>;
    using S = HeavyShape<LA, LB, LC, LD>;
    using Cf = H2Cfg<LA, LB, LC, LD, SMALL>;
    using RT = std::conditional_t<FP32, float, double>;
    using SM = std::conditional_t<RL, H2SmemRL<LA, LB, LC, LD, RT, FACT, NARROW>, H2Smem<LA, LB, LC, LD, RT, FACT, SMALL>>;
    __shared__ int rtab[RL ? 1 : S::NR];
    __shared__ double sI[S::NI];
    __shared__ SM sm;
    const int lane = threadIdx.x;
    if constexpr (!RL) heavy2_init<LA, LB, LC, LD>(rtab, lane);
    int t0 = 0, u0 = 0, v0 = 0;
    if (lane < (RL ? Cf::GL : Cf::G) * S::NHAB) hdec(lane % S::NHAB, t0, u0, v0);
    const bool doJ = (A.flags & kFlagJ) != 0, doK = (A.flags & kFlagK) != 0;
    WaveTally tally;
    UnitRange ur(C);
    for (long long u = ur.u; u < ur.end; ++u) {
        const int P = ur.owner(C, u);
        const int Q = C.kseg[P].x + (int)(u - C.unit_start[P]);
        const int2 mm = A.pmat[C.bra_off + P];
        const int4 pi = A.pinfo[C.bra_off + P];
        const double4 pg = A.pgeo[C.bra_off + P];
        const int4 qi = A.pinfo[C.ket_off + Q];
        const double4 qg = A.pgeo[C.ket_off + Q];
        if ((A.flags & kFlagScreen) && !(pg.w * qg.w * (*A.dmax) >= A.screen_tol)) continue;
        const bool small = (A.flags & kFlagMixed) && pg.w * qg.w * (*A.dmax) < A.mixed_tol;  // one quartet per wave
        if (small != FP32) continue;  // the other precision's quartet
        if constexpr (RL) heavy2rl_compute<LA, LB, LC, LD, RT, FACT, NARROW>(A, sm, sI, lane, t0, u0, v0, pi, pg, qi, qg);
        else heavy2_compute<LA, LB, LC, LD, RT, FACT, SMALL>(A, sm, rtab, sI, lane, t0, u0, v0, pi, pg, qi, qg);
        const double s = ((pi.w >> 16) ? 0.5 : 1.0) * ((qi.w >> 16) ? 0.5 : 1.0) * ((C.same && Q == P) ? 0.5 : 1.0);
        for (int t = lane; t < Dg::NTASK; t += 64) {
            int row, col;
            double val;
            if (Dg::task(t, sI, A.D + mm.x, mm.y, pi.x, pi.y, qi.x, qi.y, doJ, doK, s, row, col, val)) {
                double* tgt = (t < Dg::T2 ? A.Jt : A.Kt) + mm.x;
                unsafeAtomicAdd(&tgt[row * mm.y + col], val);
            }
        }
        tally.one(FP32, pi, qi);
        __syncthreads();  // digest reads of sI finish before the next quartet clears it
    }
    tally.flush(A, lane);
}

// ---------------------------------------------------------------------------------------------
// Launchers + dispatch table
// ---------------------------------------------------------------------------------------------
// Lane kernels loop over their work units with a grid stride: one launch per class covers every
// molecule of a batch, and the grid is capped at 64 four-wave blocks per CU (256 CUs).
constexpr long long kLaneGridCap = 256LL * 64;

template <int LA, int LB, int LC, int LD>
hipError_t launch_light(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    if (c.nunits <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((c.nunits + 3) / 4, kLaneGridCap);
    hipLaunchKernelGGL((k_jk_light<LA, LB, LC, LD, bra_in_lds(LA, LB), false>), dim3((unsigned)blocks), dim3(256), 0, s,
                       a, c);
    if (a.flags & kFlagMixed)
        hipLaunchKernelGGL((k_jk_light<LA, LB, LC, LD, bra_in_lds(LA, LB), true>), dim3((unsigned)blocks), dim3(256), 0,
                           s, a, c);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_lightc(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    if (c.nunits <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((c.nunits + 3) / 4, kLaneGridCap);
    hipLaunchKernelGGL((k_jk_lightc<LA, LB, LC, LD, bra_in_lds(LA, LB), false>), dim3((unsigned)blocks), dim3(256), 0,
                       s, a, c);
    if (a.flags & kFlagMixed)
        hipLaunchKernelGGL((k_jk_lightc<LA, LB, LC, LD, bra_in_lds(LA, LB), true>), dim3((unsigned)blocks), dim3(256),
                           0, s, a, c);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    if (c.nunits <= 0) return hipSuccess;
    const long long cap = 256LL * 64;  // 64 waves per CU of 256 CUs in flight at most
    const long long blocks = c.nunits < cap ? c.nunits : cap;
    hipLaunchKernelGGL((k_jk_heavy<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(64), 0, s, a, c);
    return hipGetLastError();
}

// Wavefront-per-quartet launches: FACT selects the factorised transforms (bit 0 bra, bit 1 ket),
// FORM the R batching (as k_jk_heavy2).
template <int LA, int LB, int LC, int LD, int FACT, int FORM>
hipError_t launch_heavy2_cfg(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    if (c.nunits <= 0) return hipSuccess;
    const long long cap = 256LL * 64;
    const long long blocks = c.nunits < cap ? c.nunits : cap;
    hipLaunchKernelGGL((k_jk_heavy2<LA, LB, LC, LD, false, FACT, FORM>), dim3((unsigned)blocks), dim3(64), 0, s,
                       a, c);
    if (a.flags & kFlagMixed)
        hipLaunchKernelGGL((k_jk_heavy2<LA, LB, LC, LD, true, FACT, FORM>), dim3((unsigned)blocks), dim3(64), 0,
                           s, a, c);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 0, 0>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2f(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 1, 0>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2s(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 0, 1>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2fs(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 1, 1>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2k(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 3, 0>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2ks(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 3, 1>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2r(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 0, 2>(a, c, s);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_heavy2rw(const BuildArgs& a, const ClassPairArgs& c, hipStream_t s)
{
    return launch_heavy2_cfg<LA, LB, LC, LD, 0, 3>(a, c, s);
}

}  // namespace qcjk
