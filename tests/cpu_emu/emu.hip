// emu.hip -- TEST HARNESS ONLY: runs the engine's __host__ __device__ quartet arithmetic on the
// CPU so the kernel math can be checked against the oracle without a GPU.  The wavefront kernel is
// emulated lane by lane, phase by phase (each phase loop over the 64 lanes stands for the code
// between two barriers).  Never linked into the product libraries.
#include <cmath>
#include <cstring>
#include <vector>

#include "mnd_quartet.h"

using namespace qcjk;

namespace {

std::vector<double> boys_table()
{
    std::vector<double> tab(kBoysRows * kBoysCols);
    for (int it = 0; it < kBoysRows; ++it) {
        const long double T = 0.1L * it, eT = expl(-T);
        for (int j = 0; j < kBoysCols; ++j) {
            long double term = 1.0L / (2 * j + 1), sum = term;
            for (int k = 1; k < 2000; ++k) {
                term *= 2.0L * T / (2 * j + 2 * k + 1);
                sum += term;
                if (term < sum * 1e-21L) break;
            }
            tab[it * kBoysCols + j] = (double)(eT * sum);
        }
    }
    return tab;
}

struct PairData {
    std::vector<double4> pp0, pp1;
    int4 info[2];
    double4 geo[2];
};

// shells: s = 0..3 with L[s], np[s], exps/coefs (folded) consumed in order, cen[3s..]
void make_pairs(const int* L, const int* np, const double* exps, const double* coefs, const double* cen, PairData& pd)
{
    int off[4], o = 0;
    for (int s = 0; s < 4; ++s) { off[s] = o; o += np[s]; }
    for (int pr = 0; pr < 2; ++pr) {
        const int A = 2 * pr, B = 2 * pr + 1;
        const int first = (int)pd.pp0.size();
        const double r2 = (cen[3 * A] - cen[3 * B]) * (cen[3 * A] - cen[3 * B]) +
                          (cen[3 * A + 1] - cen[3 * B + 1]) * (cen[3 * A + 1] - cen[3 * B + 1]) +
                          (cen[3 * A + 2] - cen[3 * B + 2]) * (cen[3 * A + 2] - cen[3 * B + 2]);
        for (int i = 0; i < np[A]; ++i)
            for (int j = 0; j < np[B]; ++j) {
                const double a = exps[off[A] + i], b = exps[off[B] + j], p = a + b;
                double P[3];
                for (int x = 0; x < 3; ++x) P[x] = (a * cen[3 * A + x] + b * cen[3 * B + x]) / p;
                pd.pp0.push_back(make_double4(P[0], P[1], P[2], p));
                pd.pp1.push_back(make_double4(P[0] - cen[3 * A], P[1] - cen[3 * A + 1], P[2] - cen[3 * A + 2],
                                              coefs[off[A] + i] * coefs[off[B] + j] * std::exp(-a * b / p * r2)));
            }
        pd.info[pr] = make_int4(0, 0, first, np[A] * np[B]);
        pd.geo[pr] = make_double4(cen[3 * A] - cen[3 * B], cen[3 * A + 1] - cen[3 * B + 1],
                                  cen[3 * A + 2] - cen[3 * B + 2], 0.0);
    }
}

template <int LA, int LB, int LC, int LD>
void run_light(const BuildArgs& A, const PairData& pd, double* out)
{
    constexpr int N = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    std::vector<double> I(N, 0.0);
    light_compute<LA, LB, LC, LD>(A, pd.info[0], pd.geo[0], pd.info[1], pd.geo[1], I.data());
    std::memcpy(out, I.data(), N * sizeof(double));
}

template <int LA, int LB, int LC, int LD>
void run_heavy(const BuildArgs& A, const PairData& pd, double* out)
{
    using S = HeavyShape<LA, LB, LC, LD>;
    auto* sm = new HeavySmem<LA, LB, LC, LD>();
    std::vector<HeavyLane<LA, LB, LC, LD>> ln(64);
    for (int l = 0; l < 64; ++l) ln[l].init(l);
    std::vector<double> iacc(64 * S::KI, 0.0), xacc(64 * S::KX);
    const int4 pi = pd.info[0], qi = pd.info[1];
    const double4 pg = pd.geo[0], qg = pd.geo[1];
    for (int ij = 0; ij < (pi.w & 0xffff); ++ij) {
        const double4 b0 = A.pp0[pi.z + ij], b1 = A.pp1[pi.z + ij];
        for (int lane = 0; lane < 3; ++lane) {
            HermiteAxis<LA, LB, double> h;
            const double pa = lane == 0 ? b1.x : (lane == 1 ? b1.y : b1.z);
            const double ab = lane == 0 ? pg.x : (lane == 1 ? pg.y : pg.z);
            h.build(0.5 / b0.w, pa, pa + ab);
            h.store(sm->Eab[lane]);
        }
        std::fill(xacc.begin(), xacc.end(), 0.0);
        for (int kl = 0; kl < (qi.w & 0xffff); ++kl) {
            const double4 k0 = A.pp0[qi.z + kl], k1 = A.pp1[qi.z + kl];
            for (int lane = 0; lane < 3; ++lane) {
                HermiteAxis<LC, LD, double> h;
                const double pc = lane == 0 ? k1.x : (lane == 1 ? k1.y : k1.z);
                const double cd = lane == 0 ? qg.x : (lane == 1 ? qg.y : qg.z);
                h.build(0.5 / k0.w, pc, pc + cd);
                h.store(sm->Ecd[lane]);
            }
            const double p = b0.w, q = k0.w, pq = p + q, alpha = p * q / pq;
            const double X = b0.x - k0.x, Y = b0.y - k0.y, Z = b0.z - k0.z;
            double F[S::L + 1];
            boys<S::L>(alpha * (X * X + Y * Y + Z * Z), A.boys, F);
            double pw[S::L + 1];
            pw[0] = 1.0;
            for (int j = 1; j <= S::L; ++j) pw[j] = pw[j - 1] * (-2.0 * alpha);
            const double lam = kTwoPi52 / (p * q * std::sqrt(pq)) * b1.w * k1.w;
            for (int j = S::L; j >= 0; --j)
                for (int lane = 0; lane < 64; ++lane)
                    heavy_r_level<LA, LB, LC, LD>(ln[lane], lane, nherm(S::L - j), pw[j] * F[j], X, Y, Z,
                                                  sm->R[(j + 1) & 1], sm->R[j & 1]);
            for (int lane = 0; lane < 64; ++lane) heavy_x_step<LA, LB, LC, LD>(ln[lane], *sm, lam, &xacc[lane * S::KX]);
        }
        for (int lane = 0; lane < 64; ++lane)
            for (int k = 0; k < S::KX; ++k) {
                const int e = lane + 64 * k;
                if (e < S::NX) sm->XI[e] = xacc[lane * S::KX + k];
            }
        for (int lane = 0; lane < 64; ++lane) heavy_i_step<LA, LB, LC, LD>(ln[lane], *sm, &iacc[lane * S::KI]);
    }
    for (int lane = 0; lane < 64; ++lane)
        for (int k = 0; k < S::KI; ++k) {
            const int e = lane + 64 * k;
            if (e < S::NI) out[e] = iacc[lane * S::KI + k];
        }
    delete sm;
}

template <int LA, int LB, int LC, int LD>
void run(const BuildArgs& A, const PairData& pd, int heavy, double* out)
{
    if (heavy) run_heavy<LA, LB, LC, LD>(A, pd, out);
    else run_light<LA, LB, LC, LD>(A, pd, out);
}

}  // namespace

extern "C" int emu_quartet(const int* L, const int* np, const double* exps, const double* coefs, const double* cen,
                           int heavy, double* out)
{
    static std::vector<double> tab = boys_table();
    PairData pd;
    make_pairs(L, np, exps, coefs, cen, pd);
    BuildArgs A{};
    A.boys = tab.data();
    A.pp0 = pd.pp0.data();
    A.pp1 = pd.pp1.data();
    const int key = ((L[0] * 3 + L[1]) * 3 + L[2]) * 3 + L[3];
#define CASE(a, b, c, d) \
    case ((a * 3 + b) * 3 + c) * 3 + d: run<a, b, c, d>(A, pd, heavy, out); return 0;
    switch (key) {
        CASE(0, 0, 0, 0) CASE(1, 0, 0, 0) CASE(1, 0, 1, 0) CASE(1, 1, 0, 0) CASE(1, 1, 1, 0) CASE(1, 1, 1, 1)
        CASE(2, 0, 0, 0) CASE(2, 0, 1, 0) CASE(2, 0, 1, 1) CASE(2, 0, 2, 0) CASE(2, 1, 0, 0) CASE(2, 1, 1, 0)
        CASE(2, 1, 1, 1) CASE(2, 1, 2, 0) CASE(2, 1, 2, 1) CASE(2, 2, 0, 0) CASE(2, 2, 1, 0) CASE(2, 2, 1, 1)
        CASE(2, 2, 2, 0) CASE(2, 2, 2, 1) CASE(2, 2, 2, 2)
    default: return -1;
    }
#undef CASE
}

// Boys values from the engine's device routine (host build), for the Boys-vs-closed-form test.
extern "C" void emu_boys(double T, int L, double* F)
{
    static std::vector<double> tab = boys_table();
    switch (L) {
    case 0: boys<0>(T, tab.data(), F); break;
    case 1: boys<1>(T, tab.data(), F); break;
    case 2: boys<2>(T, tab.data(), F); break;
    case 3: boys<3>(T, tab.data(), F); break;
    case 4: boys<4>(T, tab.data(), F); break;
    case 5: boys<5>(T, tab.data(), F); break;
    case 6: boys<6>(T, tab.data(), F); break;
    case 7: boys<7>(T, tab.data(), F); break;
    case 8: boys<8>(T, tab.data(), F); break;
    default: break;
    }
}
