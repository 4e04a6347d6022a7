// mnd_quartet.h -- contracted shell-quartet evaluation, shared by the J/K kernels, the Schwarz
// kernels and the CPU test harness.
#pragma once

#include "mnd_device.h"
#include "mnd_heavy.h"
#include "qcjk_internal.h"

namespace qcjk {

// Which contraction order the lane-per-quartet kernel uses (register pressure, measured per class).
#ifndef QCJK_TWO_PHASE_ALL
template <int LA, int LB, int LC, int LD>
constexpr bool kLightTwoPhase = LA == 2 && LB == 1 && LC == 1;
#else
template <int LA, int LB, int LC, int LD>
constexpr bool kLightTwoPhase = true;
#endif

// Lane-per-quartet: all primitive quartets of (pi,pg | qi,qg) accumulated into I (registers),
// raw (un-normalised) Cartesian integrals, layout I[((a*NB+b)*NC+c)*ND+d].
// With eab_mem (a wavefront-private LDS slot) the bra Hermite tables -- identical in every lane,
// since a wavefront shares its bra pair -- are kept there instead of in registers: every lane
// stores the same values and reads back its own stores, so no cross-lane ordering is needed.
template <int LA, int LB, int LC, int LD, bool BRA_MEM = false, typename RT = double>
QC_HD void light_compute(const BuildArgs& A, int4 pi, double4 pg, int4 qi, double4 qg, RT* I,
                         HermiteAxis<LA, LB, RT>* eab_mem = nullptr)
{
    constexpr int L = LA + LB + LC + LD;
    const int npab = pi.w & 0xffff, npcd = qi.w & 0xffff;
    for (int ij = 0; ij < npab; ++ij) {
        const double4 b0 = A.pp0[pi.z + ij], b1 = A.pp1[pi.z + ij];
        HermiteAxis<LA, LB, RT> eloc[3];
        const RT a12 = RT(0.5 / b0.w);
        eloc[0].build(a12, RT(b1.x), RT(b1.x + pg.x));
        eloc[1].build(a12, RT(b1.y), RT(b1.y + pg.y));
        eloc[2].build(a12, RT(b1.z), RT(b1.z + pg.z));
        const HermiteAxis<LA, LB, RT>* eab = eloc;
        if constexpr (BRA_MEM) {
            eloc[0].store(eab_mem[0]);
            eloc[1].store(eab_mem[1]);
            eloc[2].store(eab_mem[2]);
            eab = eab_mem;
        }
        for (int kl = 0; kl < npcd; ++kl) {
            const double4 k0 = A.pp0[qi.z + kl], k1 = A.pp1[qi.z + kl];
            HermiteAxis<LC, LD, RT> ecd[3];
            const RT c12 = RT(0.5 / k0.w);
            ecd[0].build(c12, RT(k1.x), RT(k1.x + qg.x));
            ecd[1].build(c12, RT(k1.y), RT(k1.y + qg.y));
            ecd[2].build(c12, RT(k1.z), RT(k1.z + qg.z));
            const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
            const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            RT F[L + 1];
            if constexpr (std::is_same_v<RT, double>) boys<L>(T, A.boys, F);
            else boys_mixed<L>(T, A.boys, F);
            RT R[nherm(L)];
            rtuv<L, RT>(RT(alpha), RT(X), RT(Y), RT(Z), F, R);
            const RT lam = RT(kTwoPi52 / (p * q * sqrt(pq)) * b1.w * k1.w);
            if constexpr (kLightTwoPhase<LA, LB, LC, LD>)
                prim_quartet_two_phase<LA, LB, LC, LD, RT>(eab, ecd, R, lam, I);
            else
                prim_quartet<LA, LB, LC, LD, RT>(eab, ecd, R, lam, I);
        }
    }
}

// Lane-per-quartet, one ket first-component chunk CI: I[(a*NB+b)*ND+d] for c = CI, all primitives.
template <int LA, int LB, int LC, int LD, int CI, bool BRA_MEM, typename RT = double>
QC_HD void light_compute_chunk(const BuildArgs& A, int4 pi, double4 pg, int4 qi, double4 qg, RT* I,
                               HermiteAxis<LA, LB, RT>* eab_mem)
{
    constexpr int L = LA + LB + LC + LD;
    const int npab = pi.w & 0xffff, npcd = qi.w & 0xffff;
    for (int ij = 0; ij < npab; ++ij) {
        const double4 b0 = A.pp0[pi.z + ij], b1 = A.pp1[pi.z + ij];
        HermiteAxis<LA, LB, RT> eloc[3];
        const RT a12 = RT(0.5 / b0.w);
        eloc[0].build(a12, RT(b1.x), RT(b1.x + pg.x));
        eloc[1].build(a12, RT(b1.y), RT(b1.y + pg.y));
        eloc[2].build(a12, RT(b1.z), RT(b1.z + pg.z));
        const HermiteAxis<LA, LB, RT>* eab = eloc;
        if constexpr (BRA_MEM) {
            eloc[0].store(eab_mem[0]);
            eloc[1].store(eab_mem[1]);
            eloc[2].store(eab_mem[2]);
            eab = eab_mem;
        }
        for (int kl = 0; kl < npcd; ++kl) {
            const double4 k0 = A.pp0[qi.z + kl], k1 = A.pp1[qi.z + kl];
            HermiteAxis<LC, LD, RT> ecd[3];
            const RT c12 = RT(0.5 / k0.w);
            ecd[0].build(c12, RT(k1.x), RT(k1.x + qg.x));
            ecd[1].build(c12, RT(k1.y), RT(k1.y + qg.y));
            ecd[2].build(c12, RT(k1.z), RT(k1.z + qg.z));
            const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
            const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            RT F[L + 1];
            if constexpr (std::is_same_v<RT, double>) boys<L>(T, A.boys, F);
            else boys_mixed<L>(T, A.boys, F);
            RT R[nherm(L)];
            rtuv<L, RT>(RT(alpha), RT(X), RT(Y), RT(Z), F, R);
            const RT lam = RT(kTwoPi52 / (p * q * sqrt(pq)) * b1.w * k1.w);
            prim_quartet_chunk<LA, LB, LC, LD, CI, RT>(eab, ecd, R, lam, I);
        }
    }
}

// Wavefront-cooperative quartet (see mnd_heavy.h).  SYNC() is the workgroup barrier on the
// device; the CPU harness runs the same phases lane by lane instead.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void heavy_compute(const BuildArgs& A, HeavySmem<LA, LB, LC, LD>& sm, const HeavyLane<LA, LB, LC, LD>& ln,
                              int lane, int4 pi, double4 pg, int4 qi, double4 qg)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    double iacc[S::KI];
    static_for<0, S::KI>([&](auto k) { iacc[decltype(k)::value] = 0.0; });
    const int npab = pi.w & 0xffff, npcd = qi.w & 0xffff;
    for (int ij = 0; ij < npab; ++ij) {
        const double4 b0 = A.pp0[pi.z + ij], b1 = A.pp1[pi.z + ij];
        __syncthreads();
        if (lane < 3) {
            HermiteAxis<LA, LB, double> h;
            const double pa = lane == 0 ? b1.x : (lane == 1 ? b1.y : b1.z);
            const double ab = lane == 0 ? pg.x : (lane == 1 ? pg.y : pg.z);
            h.build(0.5 / b0.w, pa, pa + ab);
            h.store(sm.Eab[lane]);
        }
        double xacc[S::KX];
        static_for<0, S::KX>([&](auto k) { xacc[decltype(k)::value] = 0.0; });
        for (int kl = 0; kl < npcd; ++kl) {
            const double4 k0 = A.pp0[qi.z + kl], k1 = A.pp1[qi.z + kl];
            __syncthreads();
            if (lane < 3) {
                HermiteAxis<LC, LD, double> h;
                const double pc = lane == 0 ? k1.x : (lane == 1 ? k1.y : k1.z);
                const double cd = lane == 0 ? qg.x : (lane == 1 ? qg.y : qg.z);
                h.build(0.5 / k0.w, pc, pc + cd);
                h.store(sm.Ecd[lane]);
            }
            const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
            const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            double F[S::L + 1];
            boys<S::L>(T, A.boys, F);
            double pw[S::L + 1];
            pw[0] = 1.0;
            static_for<1, S::L + 1>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                pw[j] = pw[j - 1] * (-2.0 * alpha);
            });
            const double lam = kTwoPi52 / (p * q * sqrt(pq)) * b1.w * k1.w;
            static_for<0, S::L + 1>([&](auto jj) {
                constexpr int j = S::L - decltype(jj)::value;
                heavy_r_level<LA, LB, LC, LD>(ln, lane, nherm(S::L - j), pw[j] * F[j], X, Y, Z, sm.R[(j + 1) & 1],
                                              sm.R[j & 1]);
                __syncthreads();
            });
            heavy_x_step<LA, LB, LC, LD>(ln, sm, lam, xacc);
        }
        __syncthreads();
        static_for<0, S::KX>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int e = lane + 64 * k;
            if (e < S::NX) sm.XI[e] = xacc[k];
        });
        __syncthreads();
        heavy_i_step<LA, LB, LC, LD>(ln, sm, iacc);
    }
    __syncthreads();
    static_for<0, S::KI>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        const int e = lane + 64 * k;
        if (e < S::NI) sm.XI[e] = iacc[k];
    });
    __syncthreads();
}

}  // namespace qcjk
