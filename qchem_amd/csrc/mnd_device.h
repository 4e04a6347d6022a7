// mnd_device.h -- McMurchie-Davidson building blocks shared by the gfx950 kernels.
//
// Everything here is __host__ __device__ so the kernel arithmetic can also be exercised on the
// CPU by the test harness (tests/cpu_emu).  All loop bounds that index register arrays are
// compile-time constants (static_for), so the arrays live in VGPRs.
//
// The mathematics restated here (not the code) follows the reference path:
//   Hermite E^{ij}_t recursion      PG_Cart_MnD/Internal/Imp/Hermite2.C:95-139
//   R^{(j)}_{tuv} recursion          Evaluators/Internal/MnD/Imp/RNLM.C:19-68
//   prim (ab|cd) Hermite double sum  PG_Cart_MnD/Imp/GaussianRF.C:446-483 (sign (-1)^{tau+nu+phi} on the ket)
//   lambda = 2 pi^{5/2} / (p q sqrt(p+q)) K_AB K_CD   GaussianRF.C:450-451
// Boys: own design -- a 0.1-spaced table of F_j(T) for T < 40 with an 8-term Taylor step
// (|dT| <= 0.05, truncation < 1e-15 relative) and downward recursion; T >= 40 uses the exact
// F_0 = sqrt(pi/T)/2 and upward recursion (erf(sqrt T) == 1 to double precision there).
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#define QC_HD __host__ __device__ __forceinline__

namespace qcjk {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi52 = 34.986836655249725;  // 2 * pi^{5/2}

constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component k of a shell with angular momentum l, reference order (m outer, l inner,
// n = L-m-l; PG_Cart/Imp/IrrepBasisSet.C:46-54).  axis 0,1,2 = n,l,m (x,y,z powers).
constexpr int comp(int l, int k, int axis)
{
    int c = 0;
    for (int m = 0; m <= l; ++m)
        for (int ll = 0; ll <= l - m; ++ll) {
            if (c == k) return axis == 0 ? l - m - ll : (axis == 1 ? ll : m);
            ++c;
        }
    return 0;
}

// Triangular index of (u, v), u + v <= M: by s = u + v, then v, for arrays sized nherm2(M).
constexpr int nherm2(int m) { return (m + 1) * (m + 2) / 2; }
constexpr int hidx2(int u, int v) { return (u + v) * (u + v + 1) / 2 + v; }

// Triangular Hermite index of (t,u,v), t+u+v = s: shells by s, then t descending, then v.
constexpr int hidx(int t, int u, int v)
{
    int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + v;
}

template <int I>
using ic = std::integral_constant<int, I>;

template <int Begin, int End, typename F>
QC_HD void static_for(F&& f)
{
    if constexpr (Begin < End) {
        f(ic<Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

// ---------------------------------------------------------------------------------------------
// Boys function
// ---------------------------------------------------------------------------------------------
constexpr int kBoysRows = 401;     // T = 0.0, 0.1, ..., 40.0
constexpr int kBoysCols = 24;      // F_0 .. F_23
constexpr double kBoysTMax = 40.0;

// F[0..L] at T.  tab: kBoysRows x kBoysCols row-major (global memory or LDS).
template <int L>
QC_HD void boys(double T, const double* __restrict__ tab, double* F)
{
    if (T < kBoysTMax) {
        int it = (int)(T * 10.0 + 0.5);
        double dt = it * 0.1 - T;
        const double* r = tab + it * kBoysCols + L;
        // sum_k r_k dt^k / k!, Horner
        double f = r[7];
        f = r[6] + dt * (1.0 / 7.0) * f;
        f = r[5] + dt * (1.0 / 6.0) * f;
        f = r[4] + dt * (1.0 / 5.0) * f;
        f = r[3] + dt * (1.0 / 4.0) * f;
        f = r[2] + dt * (1.0 / 3.0) * f;
        f = r[1] + dt * (1.0 / 2.0) * f;
        f = r[0] + dt * f;
        F[L] = f;
        if constexpr (L > 0) {
            double e = exp(-T);
            double t2 = 2.0 * T;
            static_for<0, L>([&](auto jj) {
                constexpr int j = L - 1 - decltype(jj)::value;
                F[j] = (t2 * F[j + 1] + e) * (1.0 / (2 * j + 1));
            });
        }
    } else {
        F[0] = 0.5 * sqrt(kPi / T);
        if constexpr (L > 0) {
            double e = exp(-T);
            double inv2t = 0.5 / T;
            static_for<0, L>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                F[j + 1] = ((2 * j + 1) * F[j] - e) * inv2t;
            });
        }
    }
}

// Same, FP32 arithmetic (mixed-precision quartets); T >= 40 still exact in the asymptote.
template <int L>
QC_HD void boys_f32(float T, const double* __restrict__ tab, float* F)
{
    if (T < (float)kBoysTMax) {
        int it = (int)(T * 10.0f + 0.5f);
        float dt = it * 0.1f - T;
        const double* r = tab + it * kBoysCols + L;
        float f = (float)r[5];
        f = (float)r[4] + dt * (1.0f / 5.0f) * f;
        f = (float)r[3] + dt * (1.0f / 4.0f) * f;
        f = (float)r[2] + dt * (1.0f / 3.0f) * f;
        f = (float)r[1] + dt * (1.0f / 2.0f) * f;
        f = (float)r[0] + dt * f;
        F[L] = f;
        if constexpr (L > 0) {
            float e = expf(-T);
            float t2 = 2.0f * T;
            static_for<0, L>([&](auto jj) {
                constexpr int j = L - 1 - decltype(jj)::value;
                F[j] = (t2 * F[j + 1] + e) * (1.0f / (2 * j + 1));
            });
        }
    } else {
        F[0] = 0.5f * sqrtf((float)kPi / T);
        if constexpr (L > 0) {
            float e = expf(-T);
            float inv2t = 0.5f / T;
            static_for<0, L>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                F[j + 1] = ((2 * j + 1) * F[j] - e) * inv2t;
            });
        }
    }
}

// Mixed precision: F[0..L] as float.  T >= 40 keeps the FP64 asymptote (large-argument Boys stays
// FP64); below, an FP32 Taylor step and FP32 downward recursion.
template <int L>
QC_HD void boys_mixed(double T, const double* __restrict__ tab, float* F)
{
    if (T >= kBoysTMax) {
        double Fd[L + 1];
        boys<L>(T, tab, Fd);
        static_for<0, L + 1>([&](auto j) { F[decltype(j)::value] = (float)Fd[decltype(j)::value]; });
    } else {
        boys_f32<L>((float)T, tab, F);
    }
}

// ---------------------------------------------------------------------------------------------
// Hermite expansion coefficients, one axis: E[i][j][t], i<=LA, j<=LB, t<=i+j (t=i+j+1 kept 0).
// E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + PA E^{ij}_t + (t+1) E^{ij}_{t+1}, and the same with PB for j.
// ---------------------------------------------------------------------------------------------
template <int LA, int LB, typename R>
struct HermiteAxis {
    static constexpr int NT = LA + LB + 2;
    R e[LA + 1][LB + 1][NT];

    // Element-wise copy (keeps the source in registers; a struct copy would go through scratch).
    QC_HD void store(HermiteAxis& dst) const
    {
        static_for<0, LA + 1>([&](auto ii) {
            static_for<0, LB + 1>([&](auto jj) {
                static_for<0, NT>([&](auto tt) {
                    dst.e[decltype(ii)::value][decltype(jj)::value][decltype(tt)::value] =
                        e[decltype(ii)::value][decltype(jj)::value][decltype(tt)::value];
                });
            });
        });
    }

    QC_HD void build(R a12, R PA, R PB)
    {
        static_for<0, LA + 1>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            static_for<0, NT>([&](auto tt) { e[i][0][decltype(tt)::value] = R(0); });
            if constexpr (i == 0) {
                e[0][0][0] = R(1);
            } else {
                static_for<0, i + 1>([&](auto tt) {
                    constexpr int t = decltype(tt)::value;
                    R v = PA * e[i - 1][0][t] + R(t + 1) * e[i - 1][0][t + 1];
                    if constexpr (t > 0) v += a12 * e[i - 1][0][t - 1];
                    e[i][0][t] = v;
                });
            }
        });
        static_for<1, LB + 1>([&](auto jj) {
            constexpr int j = decltype(jj)::value;
            static_for<0, LA + 1>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                static_for<0, NT>([&](auto tt) { e[i][j][decltype(tt)::value] = R(0); });
                static_for<0, i + j + 1>([&](auto tt) {
                    constexpr int t = decltype(tt)::value;
                    R v = PB * e[i][j - 1][t] + R(t + 1) * e[i][j - 1][t + 1];
                    if constexpr (t > 0) v += a12 * e[i][j - 1][t - 1];
                    e[i][j][t] = v;
                });
            });
        });
    }
};

// ---------------------------------------------------------------------------------------------
// R^{(0)}_{tuv}, t+u+v <= L, triangular storage (hidx), computed in place level by level.
// R^{(j)}_{t,u,v} = (t-1) R^{(j+1)}_{t-2,u,v} + X R^{(j+1)}_{t-1,u,v}  (t>0; else the same on u, v)
// ---------------------------------------------------------------------------------------------
// top[j] = (-2 alpha)^j F_j(T) are the R^{(j)}_{000} that start each level.
template <int L, typename Rt>
QC_HD void rtuv_top(const Rt* top, Rt X, Rt Y, Rt Z, Rt* R)
{
    R[0] = top[L];
    static_for<0, L>([&](auto jj) {
        constexpr int j = L - 1 - decltype(jj)::value;  // level being built
        static_for<0, L - j>([&](auto ss) {
            constexpr int s = L - j - decltype(ss)::value;  // descending sum
            static_for<0, s + 1>([&](auto tt) {
                constexpr int t = s - decltype(tt)::value;
                static_for<0, s - t + 1>([&](auto uu) {
                    constexpr int u = s - t - decltype(uu)::value;
                    constexpr int v = s - t - u;
                    Rt r;
                    if constexpr (t > 0) {
                        r = X * R[hidx(t - 1, u, v)];
                        if constexpr (t > 1) r += Rt(t - 1) * R[hidx(t - 2, u, v)];
                    } else if constexpr (u > 0) {
                        r = Y * R[hidx(0, u - 1, v)];
                        if constexpr (u > 1) r += Rt(u - 1) * R[hidx(0, u - 2, v)];
                    } else {
                        r = Z * R[hidx(0, 0, v - 1)];
                        if constexpr (v > 1) r += Rt(v - 1) * R[hidx(0, 0, v - 2)];
                    }
                    R[hidx(t, u, v)] = r;
                });
            });
        });
        R[0] = top[j];
    });
}

template <int L, typename Rt>
QC_HD void rtuv(Rt alpha, Rt X, Rt Y, Rt Z, const Rt* F, Rt* R)
{
    const Rt m2a = Rt(-2) * alpha;
    Rt pw = Rt(1), top[L + 1];
    static_for<0, L + 1>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        top[j] = pw * F[j];
        pw *= m2a;
    });
    rtuv_top<L, Rt>(top, X, Y, Z, R);
}

// ---------------------------------------------------------------------------------------------
// One primitive quartet, all Cartesian components, accumulated into I[((a*NB+b)*NC+c)*ND+d]:
//   I += lam * sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}
// Ket-first: x[tuv] per (c,d) component pair, then the bra sum for every (a,b).
// ---------------------------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename R>
QC_HD void prim_quartet(const HermiteAxis<LA, LB, R>* eab, const HermiteAxis<LC, LD, R>* ecd,
                        const R* Rt, R lam, R* I)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int LAB = LA + LB;
    constexpr int NHAB = nherm(LAB);
    static_for<0, NC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        static_for<0, ND>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int cx = comp(LC, c, 0), cy = comp(LC, c, 1), cz = comp(LC, c, 2);
            constexpr int dx = comp(LD, d, 0), dy = comp(LD, d, 1), dz = comp(LD, d, 2);
            R x[NHAB];
            static_for<0, LAB + 1>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                static_for<0, LAB + 1 - t>([&](auto uu) {
                    constexpr int u = decltype(uu)::value;
                    static_for<0, LAB + 1 - t - u>([&](auto vv) {
                        constexpr int v = decltype(vv)::value;
                        R acc = R(0);
                        static_for<0, cx + dx + 1>([&](auto t2) {
                            constexpr int tau = decltype(t2)::value;
                            static_for<0, cy + dy + 1>([&](auto u2) {
                                constexpr int nu = decltype(u2)::value;
                                R exy = ecd[0].e[cx][dx][tau] * ecd[1].e[cy][dy][nu];
                                static_for<0, cz + dz + 1>([&](auto v2) {
                                    constexpr int phi = decltype(v2)::value;
                                    R term = exy * ecd[2].e[cz][dz][phi] * Rt[hidx(t + tau, u + nu, v + phi)];
                                    if constexpr (((tau + nu + phi) & 1) != 0) acc -= term;
                                    else acc += term;
                                });
                            });
                        });
                        x[hidx(t, u, v)] = acc;
                    });
                });
            });
            static_for<0, NA>([&](auto aa) {
                constexpr int a = decltype(aa)::value;
                constexpr int ax = comp(LA, a, 0), ay = comp(LA, a, 1), az = comp(LA, a, 2);
                static_for<0, NB>([&](auto bb) {
                    constexpr int b = decltype(bb)::value;
                    constexpr int bx = comp(LB, b, 0), by = comp(LB, b, 1), bz = comp(LB, b, 2);
                    R acc = R(0);
                    static_for<0, ax + bx + 1>([&](auto tt) {
                        constexpr int t = decltype(tt)::value;
                        static_for<0, ay + by + 1>([&](auto uu) {
                            constexpr int u = decltype(uu)::value;
                            R exy = eab[0].e[ax][bx][t] * eab[1].e[ay][by][u];
                            static_for<0, az + bz + 1>([&](auto vv) {
                                constexpr int v = decltype(vv)::value;
                                acc += exy * eab[2].e[az][bz][v] * x[hidx(t, u, v)];
                            });
                        });
                    });
                    I[((a * NB + b) * NC + c) * ND + d] += lam * acc;
                });
            });
        });
    });
}

// Contraction restricted to ket first component CI (a "chunk"): I[(a*NB+b)*ND+d] for c = CI only.
template <int LA, int LB, int LC, int LD, int CI, typename R>
QC_HD void prim_quartet_chunk(const HermiteAxis<LA, LB, R>* eab, const HermiteAxis<LC, LD, R>* ecd,
                              const R* Rt, R lam, R* I)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), ND = ncart(LD);
    constexpr int LAB = LA + LB;
    constexpr int NHAB = nherm(LAB);
    constexpr int cx = comp(LC, CI, 0), cy = comp(LC, CI, 1), cz = comp(LC, CI, 2);
    static_for<0, ND>([&](auto dd) {
        constexpr int d = decltype(dd)::value;
        constexpr int dx = comp(LD, d, 0), dy = comp(LD, d, 1), dz = comp(LD, d, 2);
        R x[NHAB];
        static_for<0, LAB + 1>([&](auto tt) {
            constexpr int t = decltype(tt)::value;
            static_for<0, LAB + 1 - t>([&](auto uu) {
                constexpr int u = decltype(uu)::value;
                static_for<0, LAB + 1 - t - u>([&](auto vv) {
                    constexpr int v = decltype(vv)::value;
                    R acc = R(0);
                    static_for<0, cx + dx + 1>([&](auto t2) {
                        constexpr int tau = decltype(t2)::value;
                        static_for<0, cy + dy + 1>([&](auto u2) {
                            constexpr int nu = decltype(u2)::value;
                            R exy = ecd[0].e[cx][dx][tau] * ecd[1].e[cy][dy][nu];
                            static_for<0, cz + dz + 1>([&](auto v2) {
                                constexpr int phi = decltype(v2)::value;
                                R term = exy * ecd[2].e[cz][dz][phi] * Rt[hidx(t + tau, u + nu, v + phi)];
                                if constexpr (((tau + nu + phi) & 1) != 0) acc -= term;
                                else acc += term;
                            });
                        });
                    });
                    x[hidx(t, u, v)] = acc;
                });
            });
        });
        static_for<0, NA>([&](auto aa) {
            constexpr int a = decltype(aa)::value;
            constexpr int ax = comp(LA, a, 0), ay = comp(LA, a, 1), az = comp(LA, a, 2);
            static_for<0, NB>([&](auto bb) {
                constexpr int b = decltype(bb)::value;
                constexpr int bx = comp(LB, b, 0), by = comp(LB, b, 1), bz = comp(LB, b, 2);
                R acc = R(0);
                static_for<0, ax + bx + 1>([&](auto tt) {
                    constexpr int t = decltype(tt)::value;
                    static_for<0, ay + by + 1>([&](auto uu) {
                        constexpr int u = decltype(uu)::value;
                        R exy = eab[0].e[ax][bx][t] * eab[1].e[ay][by][u];
                        static_for<0, az + bz + 1>([&](auto vv) {
                            constexpr int v = decltype(vv)::value;
                            acc += exy * eab[2].e[az][bz][v] * x[hidx(t, u, v)];
                        });
                    });
                });
                I[(a * NB + b) * ND + d] += lam * acc;
            });
        });
    });
}

// Same contraction, two phases: the ket transform x[cd][tuv] for every ket component pair first
// (R is dead afterwards), then the bra transform one bra component pair at a time.  Keeps the bra
// Hermite products short-lived; used where NCD * nherm(Lab) is small.
template <int LA, int LB, int LC, int LD, typename R>
QC_HD void prim_quartet_two_phase(const HermiteAxis<LA, LB, R>* eab, const HermiteAxis<LC, LD, R>* ecd,
                                  const R* Rt, R lam, R* I)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int LAB = LA + LB;
    constexpr int NHAB = nherm(LAB);
    constexpr int NCD = NC * ND;
    R x[NCD][NHAB];
    static_for<0, NC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        static_for<0, ND>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int cx = comp(LC, c, 0), cy = comp(LC, c, 1), cz = comp(LC, c, 2);
            constexpr int dx = comp(LD, d, 0), dy = comp(LD, d, 1), dz = comp(LD, d, 2);
            static_for<0, LAB + 1>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                static_for<0, LAB + 1 - t>([&](auto uu) {
                    constexpr int u = decltype(uu)::value;
                    static_for<0, LAB + 1 - t - u>([&](auto vv) {
                        constexpr int v = decltype(vv)::value;
                        R acc = R(0);
                        static_for<0, cx + dx + 1>([&](auto t2) {
                            constexpr int tau = decltype(t2)::value;
                            static_for<0, cy + dy + 1>([&](auto u2) {
                                constexpr int nu = decltype(u2)::value;
                                R exy = ecd[0].e[cx][dx][tau] * ecd[1].e[cy][dy][nu];
                                static_for<0, cz + dz + 1>([&](auto v2) {
                                    constexpr int phi = decltype(v2)::value;
                                    R term = exy * ecd[2].e[cz][dz][phi] * Rt[hidx(t + tau, u + nu, v + phi)];
                                    if constexpr (((tau + nu + phi) & 1) != 0) acc -= term;
                                    else acc += term;
                                });
                            });
                        });
                        x[c * ND + d][hidx(t, u, v)] = lam * acc;
                    });
                });
            });
        });
    });
    static_for<0, NA>([&](auto aa) {
        constexpr int a = decltype(aa)::value;
        constexpr int ax = comp(LA, a, 0), ay = comp(LA, a, 1), az = comp(LA, a, 2);
        static_for<0, NB>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            constexpr int bx = comp(LB, b, 0), by = comp(LB, b, 1), bz = comp(LB, b, 2);
            static_for<0, ax + bx + 1>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                static_for<0, ay + by + 1>([&](auto uu) {
                    constexpr int u = decltype(uu)::value;
                    const R exy = eab[0].e[ax][bx][t] * eab[1].e[ay][by][u];
                    static_for<0, az + bz + 1>([&](auto vv) {
                        constexpr int v = decltype(vv)::value;
                        const R e = exy * eab[2].e[az][bz][v];
                        static_for<0, NCD>([&](auto k) {
                            I[(a * NB + b) * NCD + decltype(k)::value] += e * x[decltype(k)::value][hidx(t, u, v)];
                        });
                    });
                });
            });
        });
    });
}

// Algorithmic FLOPs of one primitive quartet (the metric's per-unit formula):
//   40 + 2(L+1) + 3 C(L+4,4) + 10 + 2 nh(Lab) nh(Lcd) nc(lc)nc(ld) + 2 nc(la)nc(lb) nh(Lab) nc(lc)nc(ld)
constexpr double prim_flops(int la, int lb, int lc, int ld)
{
    int L = la + lb + lc + ld;
    double c4 = (double)(L + 4) * (L + 3) * (L + 2) * (L + 1) / 24.0;
    return 40.0 + 2.0 * (L + 1) + 3.0 * c4 + 10.0 +
           2.0 * nherm(la + lb) * nherm(lc + ld) * ncart(lc) * ncart(ld) +
           2.0 * ncart(la) * ncart(lb) * nherm(la + lb) * ncart(lc) * ncart(ld);
}
// Per contracted quartet: K^4 prim flops + 12 flops per Cartesian element (J: 4, K: 8 digestion).
constexpr double digest_flops(int la, int lb, int lc, int ld)
{
    return 12.0 * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
}

}  // namespace qcjk
