// qcjk_setup.hip -- per-handle setup kernels: shell-pair precompute, Schwarz bounds, and the
// one-electron matrices (S, T, V).
//
// Pair precompute replaces the reference's Omega construction (PG_Cart_MnD/Imp/GaussianRF.C:239-252:
// p = a+b, P = (aA+bB)/p, K_AB = exp(-ab/p |AB|^2)) and its process-wide cache: the data are
// computed once per handle into HBM.  The Schwarz bound Q_AB = sqrt(max_ab (ab|ab)) has no
// reference counterpart (the reference does not screen); it orders and screens the quartet work.
// The one-electron kernel restates Overlap2C / Grad2 / Nuclear (GaussianRF.C:311-405) with the
// kinetic factor 1/2 of Kinetic.C:57.
#include "mnd_quartet.h"

namespace qcjk {

// ---------------------------------------------------------------------------------------------
// Primitive-pair data
// ---------------------------------------------------------------------------------------------
__global__ void k_pair_setup(const int4* __restrict__ pshell, const int2* __restrict__ pnp, const double* __restrict__ exps,
                             const double* __restrict__ coefs, const double* __restrict__ cen, int npairs,
                             const int* __restrict__ pp_off, double4* __restrict__ pp0, double4* __restrict__ pp1)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int4 ps = pshell[k];  // {A, B, first prim of A, first prim of B}
    const int2 np = pnp[k];     // {nprim A, nprim B}
    const double Ax = cen[3 * ps.x], Ay = cen[3 * ps.x + 1], Az = cen[3 * ps.x + 2];
    const double Bx = cen[3 * ps.y], By = cen[3 * ps.y + 1], Bz = cen[3 * ps.y + 2];
    const double r2 = (Ax - Bx) * (Ax - Bx) + (Ay - By) * (Ay - By) + (Az - Bz) * (Az - Bz);
    int o = pp_off[k];
    for (int i = 0; i < np.x; ++i) {
        const double a = exps[ps.z + i], ca = coefs[ps.z + i];
        for (int j = 0; j < np.y; ++j, ++o) {
            const double b = exps[ps.w + j], cb = coefs[ps.w + j];
            const double p = a + b, ip = 1.0 / p;
            const double Px = (a * Ax + b * Bx) * ip, Py = (a * Ay + b * By) * ip, Pz = (a * Az + b * Bz) * ip;
            pp0[o] = make_double4(Px, Py, Pz, p);
            pp1[o] = make_double4(Px - Ax, Py - Ay, Pz - Az, ca * cb * exp(-a * b * ip * r2));
        }
    }
}

hipError_t launch_pair_setup(const int4* pair_shells, const int2* pair_np, const double* exps, const double* coefs,
                             const double* centers, int npairs, double4* pp0, double4* pp1, const int* pp_off,
                             hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_setup, dim3((npairs + 255) / 256), dim3(256), 0, s, pair_shells, pair_np, exps, coefs,
                       centers, npairs, pp_off, pp0, pp1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Schwarz bounds: (P|P) per shell pair, max over the diagonal components
// ---------------------------------------------------------------------------------------------
template <int LA, int LB>
__global__ __launch_bounds__(256) void k_schwarz_light(BuildArgs A, int off, int npairs, const double* __restrict__ ns,
                                                        double* __restrict__ qraw, double* __restrict__ qnorm)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= npairs) return;
    const int4 pi = A.pinfo[off + P];
    const double4 pg = A.pgeo[off + P];
    double I[NAB * NAB];
    static_for<0, NAB * NAB>([&](auto i) { I[decltype(i)::value] = 0.0; });
    light_compute<LA, LB, LA, LB>(A, pi, pg, pi, pg, I);
    double mr = 0.0, mn = 0.0;
    static_for<0, NAB>([&](auto ab) {
        constexpr int a = decltype(ab)::value / NB, b = decltype(ab)::value % NB;
        const double v = fabs(I[decltype(ab)::value * NAB + decltype(ab)::value]);
        const double w = ns[pi.x + a] * ns[pi.y + b];
        mr = fmax(mr, v);
        mn = fmax(mn, v * w * w);
    });
    qraw[P] = sqrt(mr);
    qnorm[P] = sqrt(mn);
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz_heavy(BuildArgs A, int off, int npairs, const double* __restrict__ ns,
                                                       double* __restrict__ qraw, double* __restrict__ qnorm)
{
    using S = HeavyShape<LA, LB, LA, LB>;
    __shared__ HeavySmem<LA, LB, LA, LB> sm;
    const int lane = threadIdx.x;
    HeavyLane<LA, LB, LA, LB> ln;
    ln.init(lane);
    for (int P = blockIdx.x; P < npairs; P += gridDim.x) {
        const int4 pi = A.pinfo[off + P];
        const double4 pg = A.pgeo[off + P];
        heavy_compute<LA, LB, LA, LB>(A, sm, ln, lane, pi, pg, pi, pg);
        double mr = 0.0, mn = 0.0;
        for (int ab = lane; ab < S::NAB; ab += 64) {
            const int a = ab / S::NB, b = ab % S::NB;
            const double v = fabs(sm.XI[ab * S::NAB + ab]);
            const double w = ns[pi.x + a] * ns[pi.y + b];
            mr = fmax(mr, v);
            mn = fmax(mn, v * w * w);
        }
        for (int o = 32; o > 0; o >>= 1) {
            mr = fmax(mr, __shfl_xor(mr, o, 64));
            mn = fmax(mn, __shfl_xor(mn, o, 64));
        }
        if (lane == 0) {
            qraw[P] = sqrt(mr);
            qnorm[P] = sqrt(mn);
        }
        __syncthreads();
    }
}

hipError_t launch_schwarz(int pc, const BuildArgs& a, int off, int npairs, const double* ns, double* qraw,
                          double* qnorm, hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    const dim3 lb((npairs + 255) / 256), lt(256);
    const dim3 hb(npairs < 16384 ? npairs : 16384), ht(64);
    switch (pc) {
    case 0: hipLaunchKernelGGL((k_schwarz_light<0, 0>), lb, lt, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    case 1: hipLaunchKernelGGL((k_schwarz_light<1, 0>), lb, lt, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    case 2: hipLaunchKernelGGL((k_schwarz_heavy<1, 1>), hb, ht, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    case 3: hipLaunchKernelGGL((k_schwarz_light<2, 0>), lb, lt, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    case 4: hipLaunchKernelGGL((k_schwarz_heavy<2, 1>), hb, ht, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    case 5: hipLaunchKernelGGL((k_schwarz_heavy<2, 2>), hb, ht, 0, s, a, off, npairs, ns, qraw, qnorm); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// One-electron blocks for a shell pair (A,B), la = LA >= lb = LB; writes the block and its transpose.
// ---------------------------------------------------------------------------------------------
template <int LA, int LB>
__global__ __launch_bounds__(128) void k_one_electron(const int4* __restrict__ spairs, int nsp, const int* __restrict__ spoff,
                                                       const int* __restrict__ snp, const double* __restrict__ exps,
                                                       const double* __restrict__ coefs, const double* __restrict__ cen,
                                                       const double* __restrict__ ns, int n, int natom,
                                                       const double* __restrict__ Zn, const double* __restrict__ Rn,
                                                       const double* __restrict__ boystab, double* S, double* T, double* V)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsp) return;
    const int4 sp = spairs[k];  // {A, B, bfA, bfB}
    const double A[3] = {cen[3 * sp.x], cen[3 * sp.x + 1], cen[3 * sp.x + 2]};
    const double B[3] = {cen[3 * sp.y], cen[3 * sp.y + 1], cen[3 * sp.y + 2]};
    const double r2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    double s[NA * NB], t[NA * NB], v[NA * NB];
    static_for<0, NA * NB>([&](auto i) { s[decltype(i)::value] = t[decltype(i)::value] = v[decltype(i)::value] = 0.0; });
    for (int i = 0; i < snp[sp.x]; ++i) {
        const double a = exps[spoff[sp.x] + i], ca = coefs[spoff[sp.x] + i];
        for (int j = 0; j < snp[sp.y]; ++j) {
            const double b = exps[spoff[sp.y] + j], cb = coefs[spoff[sp.y] + j];
            const double p = a + b, ip = 1.0 / p;
            double P[3], PA[3], PB[3];
            for (int x = 0; x < 3; ++x) {
                P[x] = (a * A[x] + b * B[x]) * ip;
                PA[x] = P[x] - A[x];
                PB[x] = P[x] - B[x];
            }
            const double kab = exp(-a * b * ip * r2) * ca * cb;
            HermiteAxis<LA + 1, LB + 1, double> E[3];
            E[0].build(0.5 * ip, PA[0], PB[0]);
            E[1].build(0.5 * ip, PA[1], PB[1]);
            E[2].build(0.5 * ip, PA[2], PB[2]);
            const double pref = ip * sqrt(ip) * (kPi * sqrt(kPi)) * kab;  // (pi/p)^{3/2} K_AB c_i c_j
            static_for<0, NA>([&](auto aa) {
                constexpr int ia = decltype(aa)::value;
                constexpr int ax = comp(LA, ia, 0), ay = comp(LA, ia, 1), az = comp(LA, ia, 2);
                static_for<0, NB>([&](auto bb) {
                    constexpr int ib = decltype(bb)::value;
                    constexpr int bx = comp(LB, ib, 0), by = comp(LB, ib, 1), bz = comp(LB, ib, 2);
                    const double sx = E[0].e[ax][bx][0], sy = E[1].e[ay][by][0], sz = E[2].e[az][bz][0];
                    s[ia * NB + ib] += pref * sx * sy * sz;
                    // <-nabla^2> per axis: la lb S(la-1,lb-1) - 2 b la S(la-1,lb+1) - 2 a lb S(la+1,lb-1) + 4ab S(la+1,lb+1)
                    auto g1 = [&](const auto& e, auto ci, auto di) {
                        constexpr int c = decltype(ci)::value, d = decltype(di)::value;
                        double r = 4.0 * a * b * e.e[c + 1][d + 1][0];
                        if constexpr (c > 0) r -= 2.0 * b * c * e.e[c - 1][d + 1][0];
                        if constexpr (d > 0) r -= 2.0 * a * d * e.e[c + 1][d - 1][0];
                        if constexpr (c > 0 && d > 0) r += (double)(c * d) * e.e[c - 1][d - 1][0];
                        return r;
                    };
                    const double gx = g1(E[0], ic<ax>{}, ic<bx>{}) * sy * sz;
                    const double gy = sx * g1(E[1], ic<ay>{}, ic<by>{}) * sz;
                    const double gz = sx * sy * g1(E[2], ic<az>{}, ic<bz>{});
                    t[ia * NB + ib] += 0.5 * pref * (gx + gy + gz);
                });
            });
            for (int c = 0; c < natom; ++c) {
                const double X = P[0] - Rn[3 * c], Y = P[1] - Rn[3 * c + 1], Z = P[2] - Rn[3 * c + 2];
                double F[L + 1];
                boys<L>(p * (X * X + Y * Y + Z * Z), boystab, F);
                double R[nherm(L)];
                rtuv<L, double>(p, X, Y, Z, F, R);
                const double pv = -Zn[c] * 2.0 * kPi * ip * kab;
                static_for<0, NA>([&](auto aa) {
                    constexpr int ia = decltype(aa)::value;
                    constexpr int ax = comp(LA, ia, 0), ay = comp(LA, ia, 1), az = comp(LA, ia, 2);
                    static_for<0, NB>([&](auto bb) {
                        constexpr int ib = decltype(bb)::value;
                        constexpr int bx = comp(LB, ib, 0), by = comp(LB, ib, 1), bz = comp(LB, ib, 2);
                        double acc = 0.0;
                        static_for<0, ax + bx + 1>([&](auto tt) {
                            static_for<0, ay + by + 1>([&](auto uu) {
                                static_for<0, az + bz + 1>([&](auto vv) {
                                    constexpr int tq = decltype(tt)::value, uq = decltype(uu)::value, vq = decltype(vv)::value;
                                    acc += E[0].e[ax][bx][tq] * E[1].e[ay][by][uq] * E[2].e[az][bz][vq] * R[hidx(tq, uq, vq)];
                                });
                            });
                        });
                        v[ia * NB + ib] += pv * acc;
                    });
                });
            }
        }
    }
    static_for<0, NA>([&](auto aa) {
        constexpr int ia = decltype(aa)::value;
        static_for<0, NB>([&](auto bb) {
            constexpr int ib = decltype(bb)::value;
            const int r = sp.z + ia, c = sp.w + ib;
            const double w = ns[r] * ns[c];
            if (S) { S[r * n + c] = s[ia * NB + ib] * w; S[c * n + r] = s[ia * NB + ib] * w; }
            if (T) { T[r * n + c] = t[ia * NB + ib] * w; T[c * n + r] = t[ia * NB + ib] * w; }
            if (V) { V[r * n + c] = v[ia * NB + ib] * w; V[c * n + r] = v[ia * NB + ib] * w; }
        });
    });
}

hipError_t launch_one_electron(int la, int lb, const int4* spairs, int nsp, const int* shell_poff, const int* shell_np,
                               const double* exps, const double* coefs, const double* centers, const double* ns, int n,
                               int natom, const double* Z, const double* Rn, const double* bt, double* S, double* T,
                               double* V, hipStream_t s)
{
    if (nsp <= 0) return hipSuccess;
    const dim3 b((nsp + 127) / 128), t(128);
#define QCJK_OE(LA, LB)                                                                                               \
    hipLaunchKernelGGL((k_one_electron<LA, LB>), b, t, 0, s, spairs, nsp, shell_poff, shell_np, exps, coefs, centers, \
                       ns, n, natom, Z, Rn, bt, S, T, V)
    switch (la * 3 + lb) {
    case 0: QCJK_OE(0, 0); break;
    case 3: QCJK_OE(1, 0); break;
    case 4: QCJK_OE(1, 1); break;
    case 6: QCJK_OE(2, 0); break;
    case 7: QCJK_OE(2, 1); break;
    case 8: QCJK_OE(2, 2); break;
    default: return hipErrorInvalidValue;
    }
#undef QCJK_OE
    return hipGetLastError();
}

}  // namespace qcjk
