// mnd_heavy2.h -- wavefront-per-quartet variant with register-resident transforms ("Hermite lanes").
//
// One 64-lane wavefront owns one contracted shell quartet (like mnd_heavy.h), but the work is laid
// out so that the inner loops are fully static and run on registers, and barriers are amortised
// over many primitive quartets:
//   R batch        R^{(j)}_{tuv} for up to KB ket primitive pairs at once, level by level in LDS
//                  (one barrier per level per batch, not per primitive quartet).
//   ket transform  lane <-> (bra Hermite index (t,u,v), group): G = 64 / nherm(Lab) groups of
//                  nherm(Lab) lanes each take every G-th ket primitive pair of the batch.  Each lane
//                  gathers its window R_{t+tau,u+nu,v+phi} (tau+nu+phi <= Lcd) from LDS, then for
//                  every ket component pair cd accumulates
//                  X[cd][tuv] += lam * sum (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{...}
//                  with the ket Hermite tables in registers; the groups' partial sums meet in LDS
//                  (atomic adds).
//   bra transform  lane <-> (bra primitive pair, ket component pair cd): gathers X[cd][*] of its
//                  bra primitive pair from LDS and adds sum_tuv E^{ab}_{tuv} X[cd][tuv] into
//                  I[ab][cd] (LDS FP64 atomics) for every ab: with the bra Hermite tables in
//                  registers, or (FACT) one axis at a time (Y[u][v] = sum_t Ex X[t][u][v],
//                  Z[v] = sum_u Ey Y[u][v], I = sum_v Ez Z[v]) with the tables of the batch's bra
//                  primitive pairs in LDS: fewer registers and FLOPs, more LDS.
// FACT is a mask: bit 0 factorises the bra transform as above, bit 1 the ket transform the same way
// (Y[nu][phi] = sum_tau (-1)^tau Ex R, ...) with the batch's ket Hermite tables in LDS.
//
// Lane-R form (heavy2rl_compute): the R batch holds (bra, ket) primitive pairs of the whole X
// batch, flattened, up to KBL at a time.  Each lane evaluates the Boys function and the complete
// R^{(0)} table of one pair -- in registers for L <= 6 (at most 84 entries), in place in its own
// LDS row above -- and stores it to LDS: no barrier per R level, and every lane busy where a light
// ket has only a few ket primitive pairs per bra primitive pair.  Rows have an odd length so the
// 64 lanes' rows start in different LDS banks.
#pragma once

#include "mnd_heavy.h"

namespace qcjk {

// Batch sizes.  SMALL trades batching for occupancy: about 8 KB each for the R batch and the X
// batch, so a one-wavefront block fits ~20 KB of LDS and two blocks run per SIMD.
template <int LA, int LB, int LC, int LD, bool SMALL = false>
struct H2Cfg {
    using S = HeavyShape<LA, LB, LC, LD>;
    static constexpr int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
    static constexpr int KB0 = S::NR <= 96 ? 9 : (S::NR <= 130 ? 8 : 6);
    static constexpr int KB = SMALL ? clampi(8192 / (16 * S::NR), 3, KB0) : KB0;  // ket prim pairs per R batch
    static constexpr int XI = S::NCD * S::NHAB;
    static constexpr int IB0 = XI >= 2304 ? 1 : (2304 / XI > 9 ? 9 : 2304 / XI);
    static constexpr int IB = SMALL ? clampi(1024 / XI, 1, IB0) : IB0;  // bra prim pairs per X batch
    // ket transform: lane groups of NHAB lanes, each group on its own ket primitive pairs of the batch
    static constexpr int G = 64 / S::NHAB > KB ? KB : (64 / S::NHAB < 1 ? 1 : 64 / S::NHAB);
    // lane-R form: flattened (bra, ket) primitive pairs per R batch, about 16 KB of FP64 R tables,
    // or (narrow) about 8 KB
    static constexpr int KBL = clampi(16384 / (8 * S::NR), 16, 64);
    static constexpr int KBLN = clampi(8192 / (8 * (S::NR | 1)), 8, 64);
    static constexpr int IBN = clampi(1024 / XI, 1, IB0);
    static constexpr int GL = 64 / S::NHAB < 1 ? 1 : 64 / S::NHAB;
};

// Per-wavefront LDS of the R batch and the ket-transformed block, in the quartet's arithmetic
// type.  The R recursion table and the contracted block I (always FP64) live outside it.
template <int LA, int LB, int LC, int LD, typename RT, int FACT = 0, bool SMALL = false>
struct H2Smem {
    using S = HeavyShape<LA, LB, LC, LD>;
    using C = H2Cfg<LA, LB, LC, LD, SMALL>;
    RT R[2][C::KB][S::NR];
    RT top[C::KB][S::L + 1];  // (-2 alpha)^j F_j per batch slot
    RT geo[C::KB][4];         // P-Q (x,y,z), lambda per batch slot
    RT X[C::IB][S::NCD][S::NHAB];
    HermiteAxis<LA, LB, RT> E[(FACT & 1) ? C::IB : 1][3];   // bra Hermite tables of a batch's bra primitive pairs
    HermiteAxis<LC, LD, RT> Ek[(FACT & 2) ? C::KB : 1][3];  // ket Hermite tables of a batch's ket primitive pairs
};

// Lane-R form: R^{(0)} tables of the flattened batch, per slot its bra primitive pair (within the X
// batch), lambda and the ket Hermite inputs 1/(2q), QC.
template <int LA, int LB, int LC, int LD, typename RT, int FACT = 0, bool NARROW = false>
struct H2SmemRL {
    using S = HeavyShape<LA, LB, LC, LD>;
    using C = H2Cfg<LA, LB, LC, LD, false>;
    static constexpr int KBL = NARROW ? C::KBLN : C::KBL;  // slots per R batch
    static constexpr int IB = NARROW ? C::IBN : C::IB;     // bra primitive pairs per X batch
    static constexpr int NRP = S::NR | 1;                // odd row length
    RT R[KBL][NRP];
    double kq[KBL][4];  // 1/(2q), QC (x,y,z)
    RT lam[KBL];
    int sii[KBL];
    RT X[IB][S::NCD][S::NHAB];
    HermiteAxis<LA, LB, RT> E[(FACT & 1) ? IB : 1][3];
};

// Reference-order index of the Cartesian component (x, y, z) of a shell of angular momentum l.
constexpr int cidx(int l, int y, int z) { return z * (l + 1) - z * (z - 1) / 2 + y; }
static_assert(cidx(2, 0, 1) == 3 && comp(2, 3, 0) == 1 && comp(2, 3, 2) == 1 && cidx(2, 0, 2) == 5 &&
                  cidx(1, 1, 0) == 1 && comp(1, 1, 1) == 1,
              "component index must invert comp()");

// Build the per-entry R recursion table (once per kernel): dir | i1<<2 | i2<<10 | c2<<18.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void heavy2_init(int* rtab, int lane)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    for (int e = lane; e < S::NR; e += 64) {
        int t, u, v, dir = 0, i1 = 0, i2 = 0, c2 = 0;
        hdec(e, t, u, v);
        if (e == 0) { rtab[e] = 0; continue; }
        if (t > 0) { dir = 0; i1 = hidx(t - 1, u, v); if (t > 1) { i2 = hidx(t - 2, u, v); c2 = t - 1; } }
        else if (u > 0) { dir = 1; i1 = hidx(0, u - 1, v); if (u > 1) { i2 = hidx(0, u - 2, v); c2 = u - 1; } }
        else { dir = 2; i1 = hidx(0, 0, v - 1); if (v > 1) { i2 = hidx(0, 0, v - 2); c2 = v - 1; } }
        rtab[e] = dir | (i1 << 2) | (i2 << 10) | (c2 << 18);
    }
    __syncthreads();
}

// The lane's window rw[tau nu phi] = R_{t0+tau, u0+nu, v0+phi}, tau+nu+phi <= LCD, of one R^{(0)} table.
template <int LCD, typename RT>
__device__ __forceinline__ void ket_window(const RT* R0, int t0, int u0, int v0, RT* rw)
{
    static_for<0, LCD + 1>([&](auto ta) {
        constexpr int tau = decltype(ta)::value;
        static_for<0, LCD + 1 - tau>([&](auto na) {
            constexpr int nu = decltype(na)::value;
            static_for<0, LCD + 1 - tau - nu>([&](auto pa) {
                constexpr int phi = decltype(pa)::value;
                rw[hidx(tau, nu, phi)] = R0[hidx(t0 + tau, u0 + nu, v0 + phi)];
            });
        });
    });
}

// xacc[cd] += lam * sum (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} rw[tau nu phi], the ket Hermite
// tables built in registers from 1/(2q) and QC, QD = QC + (C - D).
template <int LC, int LD, typename RT>
__device__ __forceinline__ void ket_reg(const RT* rw, RT lam, double c12, double qcx, double qcy, double qcz,
                                        const double4& qg, RT* xacc)
{
    constexpr int NC = ncart(LC), ND = ncart(LD);
    HermiteAxis<LC, LD, RT> ecd[3];
    ecd[0].build(RT(c12), RT(qcx), RT(qcx + qg.x));
    ecd[1].build(RT(c12), RT(qcy), RT(qcy + qg.y));
    ecd[2].build(RT(c12), RT(qcz), RT(qcz + qg.z));
    static_for<0, NC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        static_for<0, ND>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int cx = comp(LC, c, 0), cy = comp(LC, c, 1), cz = comp(LC, c, 2);
            constexpr int dx = comp(LD, d, 0), dy = comp(LD, d, 1), dz = comp(LD, d, 2);
            RT acc = RT(0);
            static_for<0, cx + dx + 1>([&](auto t2) {
                constexpr int tau = decltype(t2)::value;
                static_for<0, cy + dy + 1>([&](auto u2) {
                    constexpr int nu = decltype(u2)::value;
                    const RT exy = ecd[0].e[cx][dx][tau] * ecd[1].e[cy][dy][nu];
                    static_for<0, cz + dz + 1>([&](auto v2) {
                        constexpr int phi = decltype(v2)::value;
                        const RT term = exy * ecd[2].e[cz][dz][phi] * rw[hidx(tau, nu, phi)];
                        if constexpr (((tau + nu + phi) & 1) != 0) acc -= term;
                        else acc += term;
                    });
                });
            });
            xacc[c * ND + d] += lam * acc;
        });
    });
}

// Same sum axis by axis over ket Hermite tables K[0..2] held in LDS:
// Y[nu][phi] = sum_tau (-1)^tau Ex R, then (-1)^nu Ey, then (-1)^phi Ez.
template <int LC, int LD, typename RT>
__device__ __forceinline__ void ket_fact(const RT* rw, RT lam, const HermiteAxis<LC, LD, RT>* K, RT* xacc)
{
    constexpr int ND = ncart(LD);
    const HermiteAxis<LC, LD, RT>& Kx = K[0];
    const HermiteAxis<LC, LD, RT>& Ky = K[1];
    const HermiteAxis<LC, LD, RT>& Kz = K[2];
    static_for<0, LC + 1>([&](auto cxx) {
        constexpr int cx = decltype(cxx)::value;
        static_for<0, LD + 1>([&](auto dxx) {
            constexpr int dx = decltype(dxx)::value;
            constexpr int M = (LC - cx) + (LD - dx);
            RT Y[nherm2(M)];
            static_for<0, M + 1>([&](auto nn) {
                constexpr int nu = decltype(nn)::value;
                static_for<0, M + 1 - nu>([&](auto pp) {
                    constexpr int phi = decltype(pp)::value;
                    RT y = RT(0);
                    static_for<0, cx + dx + 1>([&](auto tt) {
                        constexpr int tau = decltype(tt)::value;
                        const RT term = Kx.e[cx][dx][tau] * rw[hidx(tau, nu, phi)];
                        if constexpr ((tau & 1) != 0) y -= term;
                        else y += term;
                    });
                    Y[hidx2(nu, phi)] = y;
                });
            });
            static_for<0, LC - cx + 1>([&](auto cyy) {
                constexpr int cy = decltype(cyy)::value, cz = LC - cx - cy;
                static_for<0, LD - dx + 1>([&](auto dyy) {
                    constexpr int dy = decltype(dyy)::value, dz = LD - dx - dy;
                    RT acc = RT(0);
                    static_for<0, cz + dz + 1>([&](auto pp) {
                        constexpr int phi = decltype(pp)::value;
                        RT z = RT(0);
                        static_for<0, cy + dy + 1>([&](auto nn) {
                            constexpr int nu = decltype(nn)::value;
                            const RT term = Ky.e[cy][dy][nu] * Y[hidx2(nu, phi)];
                            if constexpr ((nu & 1) != 0) z -= term;
                            else z += term;
                        });
                        const RT term = Kz.e[cz][dz][phi] * z;
                        if constexpr ((phi & 1) != 0) acc -= term;
                        else acc += term;
                    });
                    constexpr int c = cidx(LC, cy, cz), d = cidx(LD, dy, dz);
                    xacc[c * ND + d] += lam * acc;
                });
            });
        });
    });
}

// Bra transform of an X batch of nib bra primitive pairs (sm.X complete and visible after the
// leading barrier): sI[ab][cd] += sum_tuv E^{ab}_{tuv} X[ii][cd][tuv].  SM provides X and E.
template <int LA, int LB, int LC, int LD, typename RT, int FACT, typename SM>
__device__ __forceinline__ void bra_batch(const BuildArgs& A, SM& sm, double* sI, int lane, int nib, int ij0,
                                          int4 pi, double4 pg)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    if constexpr ((FACT & 1) == 0) {
        __syncthreads();
        for (int pr = lane; pr < nib * S::NCD; pr += 64) {
            const int ii = pr / S::NCD, cd = pr - ii * S::NCD;
            const double4 b0 = A.pp0[pi.z + ij0 + ii], b1 = A.pp1[pi.z + ij0 + ii];
            HermiteAxis<LA, LB, RT> eab[3];
            const RT a12 = RT(0.5 / b0.w);
            eab[0].build(a12, RT(b1.x), RT(b1.x + pg.x));
            eab[1].build(a12, RT(b1.y), RT(b1.y + pg.y));
            eab[2].build(a12, RT(b1.z), RT(b1.z + pg.z));
            RT xw[S::NHAB];
            static_for<0, S::NHAB>([&](auto h) { xw[decltype(h)::value] = sm.X[ii][cd][decltype(h)::value]; });
            static_for<0, S::NA>([&](auto aa) {
                constexpr int a = decltype(aa)::value;
                constexpr int ax = comp(LA, a, 0), ay = comp(LA, a, 1), az = comp(LA, a, 2);
                static_for<0, S::NB>([&](auto bb) {
                    constexpr int b = decltype(bb)::value;
                    constexpr int bx = comp(LB, b, 0), by = comp(LB, b, 1), bz = comp(LB, b, 2);
                    RT acc = RT(0);
                    static_for<0, ax + bx + 1>([&](auto tt) {
                        constexpr int t = decltype(tt)::value;
                        static_for<0, ay + by + 1>([&](auto uu) {
                            constexpr int u = decltype(uu)::value;
                            const RT exy = eab[0].e[ax][bx][t] * eab[1].e[ay][by][u];
                            static_for<0, az + bz + 1>([&](auto vv) {
                                constexpr int v = decltype(vv)::value;
                                acc += exy * eab[2].e[az][bz][v] * xw[hidx(t, u, v)];
                            });
                        });
                    });
                    atomicAdd(&sI[(a * S::NB + b) * S::NCD + cd], (double)acc);
                });
            });
        }
    } else {
        for (int tk = lane; tk < nib * 3; tk += 64) {  // bra Hermite tables of the batch, one axis per lane
            const int ii = tk / 3, ax = tk - ii * 3;
            const double4 b0 = A.pp0[pi.z + ij0 + ii], b1 = A.pp1[pi.z + ij0 + ii];
            const double pa = ax == 0 ? b1.x : (ax == 1 ? b1.y : b1.z);
            const double ab = ax == 0 ? pg.x : (ax == 1 ? pg.y : pg.z);
            HermiteAxis<LA, LB, RT> e;
            e.build(RT(0.5 / b0.w), RT(pa), RT(pa + ab));
            e.store(sm.E[ii][ax]);
        }
        __syncthreads();
        for (int pr = lane; pr < nib * S::NCD; pr += 64) {
            const int ii = pr / S::NCD, cd = pr - ii * S::NCD;
            const HermiteAxis<LA, LB, RT>& Ex = sm.E[ii][0];
            const HermiteAxis<LA, LB, RT>& Ey = sm.E[ii][1];
            const HermiteAxis<LA, LB, RT>& Ez = sm.E[ii][2];
            RT xw[S::NHAB];
            static_for<0, S::NHAB>([&](auto h) { xw[decltype(h)::value] = sm.X[ii][cd][decltype(h)::value]; });
            static_for<0, LA + 1>([&](auto axx) {
                constexpr int ax = decltype(axx)::value;
                static_for<0, LB + 1>([&](auto bxx) {
                    constexpr int bx = decltype(bxx)::value;
                    constexpr int M = (LA - ax) + (LB - bx);  // u + v <= M
                    // Y[u][v] = sum_t Ex[ax][bx][t] X[t][u][v]
                    RT Y[nherm2(M)];
                    static_for<0, M + 1>([&](auto uu) {
                        constexpr int u = decltype(uu)::value;
                        static_for<0, M + 1 - u>([&](auto vv) {
                            constexpr int v = decltype(vv)::value;
                            RT y = RT(0);
                            static_for<0, ax + bx + 1>([&](auto tt) {
                                constexpr int t = decltype(tt)::value;
                                y += Ex.e[ax][bx][t] * xw[hidx(t, u, v)];
                            });
                            Y[hidx2(u, v)] = y;
                        });
                    });
                    static_for<0, LA - ax + 1>([&](auto ayy) {
                        constexpr int ay = decltype(ayy)::value, az = LA - ax - ay;
                        static_for<0, LB - bx + 1>([&](auto byy) {
                            constexpr int by = decltype(byy)::value, bz = LB - bx - by;
                            // Z[v] = sum_u Ey[ay][by][u] Y[u][v]; I = sum_v Ez[az][bz][v] Z[v]
                            RT acc = RT(0);
                            static_for<0, az + bz + 1>([&](auto vv) {
                                constexpr int v = decltype(vv)::value;
                                RT z = RT(0);
                                static_for<0, ay + by + 1>([&](auto uu) {
                                    constexpr int u = decltype(uu)::value;
                                    z += Ey.e[ay][by][u] * Y[hidx2(u, v)];
                                });
                                acc += Ez.e[az][bz][v] * z;
                            });
                            constexpr int a = cidx(LA, ay, az), b = cidx(LB, by, bz);
                            atomicAdd(&sI[(a * S::NB + b) * S::NCD + cd], (double)acc);
                        });
                    });
                });
            });
        }
    }
}

// One contracted quartet into sI (FP64, NAB x NCD).  RT = float evaluates the primitive integrals
// in FP32 (the Boys function at T >= 40 and the (-2 alpha)^j F_j and lambda products in FP64).
template <int LA, int LB, int LC, int LD, typename RT, int FACT, bool SMALL>
__device__ __forceinline__ void heavy2_compute(const BuildArgs& A, H2Smem<LA, LB, LC, LD, RT, FACT, SMALL>& sm,
                                               const int* __restrict__ rtab, double* sI, int lane, int t0, int u0,
                                               int v0, int4 pi, double4 pg, int4 qi, double4 qg)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    using C = H2Cfg<LA, LB, LC, LD, SMALL>;
    constexpr int NHCD = nherm(S::LCD);
    for (int e = lane; e < S::NI; e += 64) sI[e] = 0.0;
    const int npab = pi.w & 0xffff, npcd = qi.w & 0xffff;
    const bool xl = lane < C::G * S::NHAB;
    const int grp = lane / S::NHAB, hl = lane - grp * S::NHAB;  // ket-transform group, Hermite index
    for (int ij0 = 0; ij0 < npab; ij0 += C::IB) {
        const int nib = npab - ij0 < C::IB ? npab - ij0 : C::IB;
        if constexpr (C::G > 1) {
            RT* X = &sm.X[0][0][0];
            for (int e = lane; e < C::IB * S::NCD * S::NHAB; e += 64) X[e] = RT(0);  // seen after the next barrier
        }
        for (int ii = 0; ii < nib; ++ii) {
            const double4 b0 = A.pp0[pi.z + ij0 + ii], b1 = A.pp1[pi.z + ij0 + ii];
            RT xacc[S::NCD];
            static_for<0, S::NCD>([&](auto k) { xacc[decltype(k)::value] = RT(0); });
            for (int kl0 = 0; kl0 < npcd; kl0 += C::KB) {
                const int nb = npcd - kl0 < C::KB ? npcd - kl0 : C::KB;
                __syncthreads();  // readers of the previous batch (R, top, geo) are done
                if (lane < nb) {
                    const double4 k0 = A.pp0[qi.z + kl0 + lane], k1 = A.pp1[qi.z + kl0 + lane];
                    const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
                    const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
                    const double T = alpha * (X * X + Y * Y + Z * Z);
                    RT F[S::L + 1];
                    if constexpr (std::is_same_v<RT, double>) boys<S::L>(T, A.boys, F);
                    else boys_mixed<S::L>(T, A.boys, F);
                    double pw = 1.0;
                    static_for<0, S::L + 1>([&](auto jj) {
                        constexpr int j = decltype(jj)::value;
                        sm.top[lane][j] = RT(pw * (double)F[j]);
                        pw *= -2.0 * alpha;
                    });
                    sm.geo[lane][0] = RT(X);
                    sm.geo[lane][1] = RT(Y);
                    sm.geo[lane][2] = RT(Z);
                    sm.geo[lane][3] = RT(kTwoPi52 / (p * q * sqrt(pq)) * b1.w * k1.w);
                }
                if constexpr ((FACT & 2) != 0) {
                    for (int tk = lane; tk < nb * 3; tk += 64) {  // ket Hermite tables, one axis per lane
                        const int s = tk / 3, ax = tk - s * 3;
                        const double4 k0 = A.pp0[qi.z + kl0 + s], k1 = A.pp1[qi.z + kl0 + s];
                        const double qc = ax == 0 ? k1.x : (ax == 1 ? k1.y : k1.z);
                        const double cd = ax == 0 ? qg.x : (ax == 1 ? qg.y : qg.z);
                        HermiteAxis<LC, LD, RT> e;
                        e.build(RT(0.5 / k0.w), RT(qc), RT(qc + cd));
                        e.store(sm.Ek[s][ax]);
                    }
                }
                __syncthreads();
                static_for<0, S::L + 1>([&](auto jj) {
                    constexpr int j = S::L - decltype(jj)::value;
                    constexpr int nv = nherm(S::L - j);
                    for (int idx = lane; idx < nb * nv; idx += 64) {
                        const int s = idx / nv, e = idx - s * nv;
                        RT* cur = sm.R[j & 1][s];
                        if (e == 0) {
                            cur[0] = sm.top[s][j];
                        } else {
                            const RT* prev = sm.R[(j + 1) & 1][s];
                            const int d = rtab[e];
                            const int dir = d & 3, i1 = (d >> 2) & 255, i2 = (d >> 10) & 255, c2 = (d >> 18) & 15;
                            cur[e] = sm.geo[s][dir] * prev[i1] + (RT)c2 * prev[i2];
                        }
                    }
                    __syncthreads();
                });
                if (xl) {
                    for (int s = grp; s < nb; s += C::G) {
                        RT rw[NHCD];
                        ket_window<S::LCD, RT>(sm.R[0][s], t0, u0, v0, rw);
                        const RT lam = sm.geo[s][3];
                        if constexpr ((FACT & 2) == 0) {
                            const double4 k0 = A.pp0[qi.z + kl0 + s], k1 = A.pp1[qi.z + kl0 + s];
                            ket_reg<LC, LD, RT>(rw, lam, 0.5 / k0.w, k1.x, k1.y, k1.z, qg, xacc);
                        } else {
                            ket_fact<LC, LD, RT>(rw, lam, sm.Ek[s], xacc);
                        }
                    }
                }
            }
            if (xl) {
                if constexpr (C::G > 1)
                    static_for<0, S::NCD>([&](auto k) { atomicAdd(&sm.X[ii][decltype(k)::value][hl], xacc[decltype(k)::value]); });
                else
                    static_for<0, S::NCD>([&](auto k) { sm.X[ii][decltype(k)::value][hl] = xacc[decltype(k)::value]; });
            }
        }
        bra_batch<LA, LB, LC, LD, RT, FACT>(A, sm, sI, lane, nib, ij0, pi, pg);
        __syncthreads();
    }
}

// Lane-R form of heavy2_compute (see the file header); FACT bit 0 as there.
template <int LA, int LB, int LC, int LD, typename RT, int FACT, bool NARROW>
__device__ __forceinline__ void heavy2rl_compute(const BuildArgs& A, H2SmemRL<LA, LB, LC, LD, RT, FACT, NARROW>& sm,
                                                 double* sI, int lane, int t0, int u0, int v0, int4 pi, double4 pg,
                                                 int4 qi, double4 qg)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    using C = H2Cfg<LA, LB, LC, LD, false>;
    using SMT = H2SmemRL<LA, LB, LC, LD, RT, FACT, NARROW>;
    constexpr int NHCD = nherm(S::LCD);
    for (int e = lane; e < S::NI; e += 64) sI[e] = 0.0;
    const int npab = pi.w & 0xffff, npcd = qi.w & 0xffff;
    const bool xl = lane < C::GL * S::NHAB;
    const int grp = lane / S::NHAB, hl = lane - grp * S::NHAB;  // ket-transform group, Hermite index
    for (int ij0 = 0; ij0 < npab; ij0 += SMT::IB) {
        const int nib = npab - ij0 < SMT::IB ? npab - ij0 : SMT::IB;
        {
            RT* X = &sm.X[0][0][0];
            for (int e = lane; e < SMT::IB * S::NCD * S::NHAB; e += 64) X[e] = RT(0);  // seen after the next barrier
        }
        const int nf = nib * npcd;  // (bra, ket) primitive pairs of this X batch, bra-major
        constexpr int KBL = SMT::KBL;
        for (int f0 = 0; f0 < nf; f0 += KBL) {
            const int nb = nf - f0 < KBL ? nf - f0 : KBL;
            __syncthreads();  // readers of the previous R batch are done
            if (lane < nb) {
                const int f = f0 + lane, ii = f / npcd, kl = f - ii * npcd;
                const double4 b0 = A.pp0[pi.z + ij0 + ii], b1 = A.pp1[pi.z + ij0 + ii];
                const double4 k0 = A.pp0[qi.z + kl], k1 = A.pp1[qi.z + kl];
                const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
                const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
                const double T = alpha * (X * X + Y * Y + Z * Z);
                RT F[S::L + 1], top[S::L + 1];
                if constexpr (std::is_same_v<RT, double>) boys<S::L>(T, A.boys, F);
                else boys_mixed<S::L>(T, A.boys, F);
                double pw = 1.0;
                static_for<0, S::L + 1>([&](auto jj) {
                    constexpr int j = decltype(jj)::value;
                    top[j] = RT(pw * (double)F[j]);
                    pw *= -2.0 * alpha;
                });
                if constexpr (S::NR <= 84) {
                    RT R[S::NR];
                    rtuv_top<S::L, RT>(top, RT(X), RT(Y), RT(Z), R);
                    static_for<0, S::NR>([&](auto e) { sm.R[lane][decltype(e)::value] = R[decltype(e)::value]; });
                } else {
                    rtuv_top<S::L, RT>(top, RT(X), RT(Y), RT(Z), &sm.R[lane][0]);
                }
                sm.kq[lane][0] = 0.5 / q;
                sm.kq[lane][1] = k1.x;
                sm.kq[lane][2] = k1.y;
                sm.kq[lane][3] = k1.z;
                sm.lam[lane] = RT(kTwoPi52 / (p * q * sqrt(pq)) * b1.w * k1.w);
                sm.sii[lane] = ii;
            }
            __syncthreads();
            if (xl) {
                RT xacc[S::NCD];
                int cur = -1;
                auto flush = [&]() {
                    if constexpr (C::GL > 1)
                        static_for<0, S::NCD>([&](auto k) { atomicAdd(&sm.X[cur][decltype(k)::value][hl], xacc[decltype(k)::value]); });
                    else
                        static_for<0, S::NCD>([&](auto k) { sm.X[cur][decltype(k)::value][hl] += xacc[decltype(k)::value]; });
                };
                for (int s = grp; s < nb; s += C::GL) {
                    const int ii = sm.sii[s];
                    if (ii != cur) {  // slots are bra-major: a lane's bra primitive pair changes rarely
                        if (cur >= 0) flush();
                        cur = ii;
                        static_for<0, S::NCD>([&](auto k) { xacc[decltype(k)::value] = RT(0); });
                    }
                    RT rw[NHCD];
                    ket_window<S::LCD, RT>(sm.R[s], t0, u0, v0, rw);
                    ket_reg<LC, LD, RT>(rw, sm.lam[s], sm.kq[s][0], sm.kq[s][1], sm.kq[s][2], sm.kq[s][3], qg, xacc);
                }
                if (cur >= 0) flush();
            }
        }
        bra_batch<LA, LB, LC, LD, RT, FACT>(A, sm, sI, lane, nib, ij0, pi, pg);
        __syncthreads();
    }
}

}  // namespace qcjk
