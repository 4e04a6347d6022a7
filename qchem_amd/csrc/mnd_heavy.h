// mnd_heavy.h -- wavefront-cooperative evaluation of one contracted shell quartet (gfx950).
//
// Used for the angular-momentum classes whose Cartesian block does not fit one lane's registers
// (e.g. (dp|dp) has 324 integrals, (dd|dd) 1296).  One 64-lane wavefront owns one quartet; its
// working set lives in LDS:
//   sR   : R^{(j)}_{tuv} ping-pong levels (triangular, nherm(L) each)   -- built level by level,
//          one barrier per level, every lane producing entries lane, lane+64, ...
//   sEab : per-axis Hermite tables of the current bra primitive pair
//   sEcd : per-axis Hermite tables of the current ket primitive pair
//   sXI  : X[cd][tuv] = sum_kl lam * sum (-1)^.. E^{cd} R   (ket transform, summed over ket primitives)
//          then, after the last bra primitive pair, the finished Cartesian block I[ab][cd]
// Each lane keeps its share of X and I in registers (entries lane + 64k), so contraction over
// primitives costs no LDS traffic; the bra transform runs once per bra primitive pair, not per
// primitive quartet.
//
// The phases are written as functions of (lane, shared block) so the CPU test harness can run
// them lane by lane with the same arithmetic (tests/cpu_emu).
#pragma once

#include "mnd_device.h"

namespace qcjk {

// (t,u,v) of triangular Hermite index e (inverse of hidx).
QC_HD void hdec(int e, int& t, int& u, int& v)
{
    int s = 0;
    while ((s + 1) * (s + 2) * (s + 3) / 6 <= e) ++s;
    int r = e - s * (s + 1) * (s + 2) / 6;
    // within shell s: block for t has (s-t+1) entries, t descending
    int tt = s, cnt = 0;
    while (cnt + (s - tt + 1) <= r) { cnt += s - tt + 1; --tt; }
    t = tt;
    v = r - cnt;
    u = s - t - v;
}

template <int LA, int LB, int LC, int LD>
struct HeavyShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NR = nherm(L), NHAB = nherm(LAB);
    static constexpr int NX = NCD * NHAB, NI = NAB * NCD;
    static constexpr int KR = (NR + 63) / 64, KX = (NX + 63) / 64, KI = (NI + 63) / 64;
    static constexpr int NXI = NX > NI ? NX : NI;
};

template <int LA, int LB, int LC, int LD>
struct HeavySmem {
    using S = HeavyShape<LA, LB, LC, LD>;
    double R[2][S::NR];
    HermiteAxis<LA, LB, double> Eab[3];
    HermiteAxis<LC, LD, double> Ecd[3];
    double XI[S::NXI];
};

// Per-lane decoded work (computed once per kernel).
template <int LA, int LB, int LC, int LD>
struct HeavyLane {
    using S = HeavyShape<LA, LB, LC, LD>;
    int rdec[S::KR];  // dir | i1<<2 | i2<<10 | c2<<18 ; -1 invalid ; entry 0 -> 0
    int xdec[S::KX];  // cx|cy<<2|cz<<4|dx<<6|dy<<8|dz<<10|t<<12|u<<16|v<<20 ; -1 invalid
    int idec[S::KI];  // ax|ay<<2|az<<4|bx<<6|by<<8|bz<<10|cd<<12 ; -1 invalid

    QC_HD void init(int lane)
    {
        static_for<0, S::KR>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int e = lane + 64 * k;
            int t, u, v, dir = 0, i1 = 0, i2 = 0, c2 = 0;
            if (e >= S::NR) { rdec[k] = -1; return; }
            if (e == 0) { rdec[k] = 0; return; }
            hdec(e, t, u, v);
            if (t > 0) { dir = 0; i1 = hidx(t - 1, u, v); if (t > 1) { i2 = hidx(t - 2, u, v); c2 = t - 1; } }
            else if (u > 0) { dir = 1; i1 = hidx(0, u - 1, v); if (u > 1) { i2 = hidx(0, u - 2, v); c2 = u - 1; } }
            else { dir = 2; i1 = hidx(0, 0, v - 1); if (v > 1) { i2 = hidx(0, 0, v - 2); c2 = v - 1; } }
            rdec[k] = dir | (i1 << 2) | (i2 << 10) | (c2 << 18);
        });
        static_for<0, S::KX>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int e = lane + 64 * k;
            if (e >= S::NX) { xdec[k] = -1; return; }
            int cd = e / S::NHAB, h = e % S::NHAB, t, u, v;
            int c = cd / S::ND, d = cd % S::ND;
            hdec(h, t, u, v);
            xdec[k] = comp(LC, c, 0) | (comp(LC, c, 1) << 2) | (comp(LC, c, 2) << 4) | (comp(LD, d, 0) << 6) |
                      (comp(LD, d, 1) << 8) | (comp(LD, d, 2) << 10) | (t << 12) | (u << 16) | (v << 20);
        });
        static_for<0, S::KI>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int e = lane + 64 * k;
            if (e >= S::NI) { idec[k] = -1; return; }
            int ab = e / S::NCD, cd = e % S::NCD;
            int a = ab / S::NB, b = ab % S::NB;
            idec[k] = comp(LA, a, 0) | (comp(LA, a, 1) << 2) | (comp(LA, a, 2) << 4) | (comp(LB, b, 0) << 6) |
                      (comp(LB, b, 1) << 8) | (comp(LB, b, 2) << 10) | (cd << 12);
        });
    }
};

// ---- phase helpers (one lane each) ----------------------------------------------------------

// Level j of the R recursion: cur[e] from prev[] for this lane's entries e < nherm(L-j).
template <int LA, int LB, int LC, int LD>
QC_HD void heavy_r_level(const HeavyLane<LA, LB, LC, LD>& ln, int lane, int nvalid, double top, double X, double Y,
                         double Z, const double* prev, double* cur)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    static_for<0, S::KR>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        const int e = lane + 64 * k;
        if (e < nvalid) {
            const int d = ln.rdec[k];
            if (e == 0) {
                cur[0] = top;
            } else {
                const int dir = d & 3, i1 = (d >> 2) & 255, i2 = (d >> 10) & 255, c2 = (d >> 18) & 15;
                const double x = dir == 0 ? X : (dir == 1 ? Y : Z);
                cur[e] = x * prev[i1] + (double)c2 * prev[i2];
            }
        }
    });
}

// Ket transform for this lane's X entries, accumulated into xacc with weight lam.
template <int LA, int LB, int LC, int LD>
QC_HD void heavy_x_step(const HeavyLane<LA, LB, LC, LD>& ln, const HeavySmem<LA, LB, LC, LD>& sm, double lam,
                        double* xacc)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    const double* R0 = sm.R[0];
    static_for<0, S::KX>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        const int d = ln.xdec[k];
        if (d < 0) return;
        int cx = d & 3, cy = (d >> 2) & 3, cz = (d >> 4) & 3, dx = (d >> 6) & 3, dy = (d >> 8) & 3, dz = (d >> 10) & 3;
        int t = (d >> 12) & 15, u = (d >> 16) & 15, v = (d >> 20) & 15;
        double acc = 0.0;
        for (int tau = 0; tau <= cx + dx; ++tau) {
            double ex = sm.Ecd[0].e[cx][dx][tau];
            for (int nu = 0; nu <= cy + dy; ++nu) {
                double exy = ex * sm.Ecd[1].e[cy][dy][nu];
                for (int phi = 0; phi <= cz + dz; ++phi) {
                    double term = exy * sm.Ecd[2].e[cz][dz][phi] * R0[hidx(t + tau, u + nu, v + phi)];
                    acc += ((tau + nu + phi) & 1) ? -term : term;
                }
            }
        }
        xacc[k] += lam * acc;
    });
}

// Bra transform for this lane's I entries from X (in sm.XI, layout [cd][tuv]).
template <int LA, int LB, int LC, int LD>
QC_HD void heavy_i_step(const HeavyLane<LA, LB, LC, LD>& ln, const HeavySmem<LA, LB, LC, LD>& sm, double* iacc)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    static_for<0, S::KI>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        const int d = ln.idec[k];
        if (d < 0) return;
        int ax = d & 3, ay = (d >> 2) & 3, az = (d >> 4) & 3, bx = (d >> 6) & 3, by = (d >> 8) & 3, bz = (d >> 10) & 3;
        int cd = (d >> 12) & 63;
        const double* X = sm.XI + cd * S::NHAB;
        double acc = 0.0;
        for (int t = 0; t <= ax + bx; ++t) {
            double ex = sm.Eab[0].e[ax][bx][t];
            for (int u = 0; u <= ay + by; ++u) {
                double exy = ex * sm.Eab[1].e[ay][by][u];
                for (int v = 0; v <= az + bz; ++v) acc += exy * sm.Eab[2].e[az][bz][v] * X[hidx(t, u, v)];
            }
        }
        iacc[k] += acc;
    });
}

// Digest task list: J_ab, J_cd, K_ac, K_bc, K_ad, K_bd.  Task t of NTASK; returns false if skipped.
template <int LA, int LB, int LC, int LD>
struct HeavyDigest {
    using S = HeavyShape<LA, LB, LC, LD>;
    static constexpr int T0 = 0, T1 = S::NAB, T2 = T1 + S::NCD, T3 = T2 + S::NA * S::NC, T4 = T3 + S::NB * S::NC,
                         T5 = T4 + S::NA * S::ND, T6 = T5 + S::NB * S::ND;
    static constexpr int NTASK = T6;

    // Computes the value and target (row, col) of task t, or returns false if J/K not requested.
    QC_HD static bool task(int t, const double* I, const double* D, int n, int oa, int ob, int oc, int od,
                           bool doJ, bool doK, double s, int& row, int& col, double& val)
    {
        constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NCD = S::NCD;
        double v = 0.0;
        if (t < T1) {
            if (!doJ) return false;
            int a = t / NB, b = t % NB;
            for (int c = 0; c < NC; ++c)
                for (int d = 0; d < ND; ++d) v += I[t * NCD + c * ND + d] * D[(oc + c) * n + od + d];
            row = oa + a; col = ob + b; val = 2.0 * s * v;
        } else if (t < T2) {
            if (!doJ) return false;
            int cd = t - T1, c = cd / ND, d = cd % ND;
            for (int a = 0; a < NA; ++a)
                for (int b = 0; b < NB; ++b) v += I[(a * NB + b) * NCD + cd] * D[(oa + a) * n + ob + b];
            row = oc + c; col = od + d; val = 2.0 * s * v;
        } else if (t < T3) {
            if (!doK) return false;
            int x = t - T2, a = x / NC, c = x % NC;
            for (int b = 0; b < NB; ++b)
                for (int d = 0; d < ND; ++d) v += I[(a * NB + b) * NCD + c * ND + d] * D[(ob + b) * n + od + d];
            row = oa + a; col = oc + c; val = s * v;
        } else if (t < T4) {
            if (!doK) return false;
            int x = t - T3, b = x / NC, c = x % NC;
            for (int a = 0; a < NA; ++a)
                for (int d = 0; d < ND; ++d) v += I[(a * NB + b) * NCD + c * ND + d] * D[(oa + a) * n + od + d];
            row = ob + b; col = oc + c; val = s * v;
        } else if (t < T5) {
            if (!doK) return false;
            int x = t - T4, a = x / ND, d = x % ND;
            for (int b = 0; b < NB; ++b)
                for (int c = 0; c < NC; ++c) v += I[(a * NB + b) * NCD + c * ND + d] * D[(ob + b) * n + oc + c];
            row = oa + a; col = od + d; val = s * v;
        } else {
            if (!doK) return false;
            int x = t - T5, b = x / ND, d = x % ND;
            for (int a = 0; a < NA; ++a)
                for (int c = 0; c < NC; ++c) v += I[(a * NB + b) * NCD + c * ND + d] * D[(oa + a) * n + oc + c];
            row = ob + b; col = od + d; val = s * v;
        }
        return true;
    }
};

}  // namespace qcjk
