// qcjk_internal.h -- device-side data layout shared by the kernels and the C-ABI layer.
//
// HBM layout (all resident for the life of a handle; nothing O(N^4) is stored):
//   molecules      : a handle holds one molecule or a batch; molecule m's matrices are an
//                    n_m x n_m row-major block at element offset base_m of one packed array.
//   shell pairs    : one global array, grouped by pair class (la>=lb: ss, ps, pp, ds, dp, dd),
//                    within a class by molecule, each molecule's run sorted by descending Schwarz
//                    bound.  Per pair an int4 {bf offset of the first shell in its molecule, the
//                    same for the second, first primitive pair, nprim pairs | same shell flag << 16},
//                    a double4 {A-B (x,y,z), Q_raw} and an int2 {base_m, n_m}.
//   primitive pairs: two double4 per primitive pair: {P (x,y,z), p} and {P-A (x,y,z), c_i c_j K_AB}.
//   density / Fock : D' = ns (x) ns o D, the unsymmetrised accumulators Jt, Kt (packed blocks).
//   Boys table     : 401 x 24 FP64 (77 KB).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qcjk {

constexpr int kNumPairClasses = 6;  // (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
constexpr int kPairClassL[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
constexpr int pair_class(int la, int lb) { return la == 0 ? 0 : (la == 1 ? 1 + lb : 3 + lb); }

enum : unsigned { kFlagJ = 1u, kFlagK = 2u, kFlagMixed = 4u, kFlagScreen = 8u };

struct BuildArgs {
    const double* D;       // D', ns-prescaled, one n_m x n_m block per molecule (packed)
    double* Jt;            // J accumulator (same layout), symmetrised later
    double* Kt;            // K accumulator
    const int2* pmat;      // per shell pair: {element offset of its molecule's block, n_m}
    unsigned flags;
    const double* boys;    // kBoysRows x kBoysCols
    const double4* pp0;    // {Px,Py,Pz,p}
    const double4* pp1;    // {PAx,PAy,PAz,pref}
    const int4* pinfo;     // per shell pair
    const double4* pgeo;   // per shell pair {ABx,ABy,ABz,Qraw}
    const double* dmax;    // device scalar max|D'|
    double screen_tol;
    double mixed_tol;
    unsigned long long* counters;  // [0] quartets computed, [1] fp32 quartets, [2] prim quartets
};

struct ClassPairArgs {
    int bra_off, nbra;     // bra class range in the pair array
    int ket_off, nket;     // ket class range
    int same;              // bra class == ket class -> ket index <= bra index
    const long long* unit_start;  // [nbra+1] prefix sums of work units per bra pair
    const int2* kseg;             // [nbra] {first ket pair of the bra's molecule, surviving ket pairs}
    long long nunits;
    int tile;              // ket pairs per unit (64 for the lane-per-quartet kernels, 1 otherwise)
};

// Launch one class-pair kernel.  Returns hipError_t.
typedef hipError_t (*ClassLauncher)(const BuildArgs&, const ClassPairArgs&, hipStream_t);

// Kernel variants of one class: lane per quartet (tile 64, absent for large classes) and the
// wavefront-per-quartet kernels (tile 1, always present).  tile 64 marks the classes whose untuned
// default is the lane-per-quartet kernel.
struct ClassKernelInfo {
    int la, lb, lc, ld;
    int tile;
    ClassLauncher light;   // may be null
    ClassLauncher heavy;   // LDS-staged transforms (mnd_heavy.h)
    ClassLauncher heavy2;  // register-resident transforms (mnd_heavy2.h)
    ClassLauncher lightc;  // lane per quartet, chunked over the ket's first component (may be null)
    ClassLauncher heavy2f; // heavy2 with the sum-factorised bra transform
    ClassLauncher heavy2s, heavy2fs;  // heavy2 / heavy2f with small LDS batches
    ClassLauncher heavy2k, heavy2ks;  // heavy2 with both transforms factorised, big / small batches
    ClassLauncher heavy2r, heavy2rw;  // heavy2 in lane-R form, R batches of ~16 KB / ~8 KB at <= 256 registers
};

// Table of the 21 bra-class >= ket-class combinations (qcjk_eri.hip).
const ClassKernelInfo* class_kernel(int bra_class, int ket_class);

// Setup kernels (qcjk_setup.hip).
hipError_t launch_pair_setup(const int4* pair_shells /* {A,B,poffA,poffB} */, const int2* pair_np,
                             const double* exps, const double* coefs, const double* centers,
                             int npairs, double4* pp0, double4* pp1, const int* pp_off, hipStream_t s);
hipError_t launch_schwarz(int pair_class, const BuildArgs& a, int off, int npairs, const double* ns,
                          double* qraw, double* qnorm, hipStream_t s);
hipError_t launch_one_electron(int la, int lb, const int4* spairs /* {A,B,bfA,bfB} */, int nsp,
                               const int* shell_poff, const int* shell_np, const double* exps,
                               const double* coefs, const double* centers, const double* ns, int n,
                               int natom, const double* Z, const double* Rn, const double* boys, double* S,
                               double* T, double* V, hipStream_t s);

}  // namespace qcjk
