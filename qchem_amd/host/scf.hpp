// scf.hpp -- restricted Hartree-Fock driver that reaches J and K only through the HF faces.
//
// Restates the reference recipe for the molecular HF path:
//   Hamiltonian   F = 1/2<-nabla^2> + V_en + J[D] - 1/2 K[D]     Hamiltonian/Internal/Imp/Hamiltonians.C:36-43
//   energy        E = Tr(D h) + 1/2 Tr(D J) - 1/4 Tr(D K) + E_nn  Vee.C:41-47, Vxc.C:36-41, IonIon.C:58-67
//   seed          core guess (diagonalise h) or an explicit density  ChargeDensity/Imp/Seed.C:64-77,
//                                                                    SCFIterator.C:153-179
//   DIIS          F' = V^T F V, error [F',D'], bordered Pulay solve, prune while svmin < SVTol,
//                 extrapolate only when ||[F',D']|| < FDMax = Z^2 0.1/32   SCFAcceleratorDIIS.C:35-212,
//                                                                    Math/DIIS.C:24-51, Calculation.C:184
//   loop          Fock from D, accelerator, diagonalise, aufbau fill, D = 2 sum_occ C C^T
//                                                                    SCFIterator.C:240-362, LoopDriver.C:56-70
//   mixing        D_in <- r D_out + (1-r) D_in, adaptive r (re-damp to r/4 when [F,D] rises)
//                                                                    DensityMixer.C:104-139
#pragma once

#include <vector>

#include "basis.hpp"
#include "hf_faces.hpp"
#include "linalg.hpp"

namespace qchem_amd {

struct Atom {
    double Z;
    double R[3];
};

struct SCFParams {
    int max_iter = 100;
    double e_tol = 1e-8;     // |E_k - E_{k-1}| (Eh)
    double fd_tol = 1e-5;    // ||[F',D']||  (MinFD, SCFParams.C:31)
    int nproj = 4;           // DIISParams Nproj
    double fd_max = -1.0;    // < 0: Z^2 0.1/32
    double sv_tol = 5e-9;    // DIISParams SVTol
    bool verbose = false;
    int guess = 0;           // 0 core Hamiltonian, 1 SAD (used when no explicit seed is given)
    int incremental = 0;     // > 0: J/K from the density change, full rebuild every `incremental` iterations
};

struct SCFResult {
    double energy = 0, e_nuc = 0, e_one = 0, e_coul = 0, e_exch = 0;
    int iterations = 0;
    bool converged = false;
    Mat D;
    std::vector<double> eps;
    std::vector<double> energy_history;  // E after each Fock build
    std::vector<double> time_history;    // seconds since start at each E
    double t_total = 0, t_jk = 0, t_diag = 0, t_setup = 0;
    long long quartets = 0, quartets_fp32 = 0;  // over all Fock builds of this run
};

double nuclear_repulsion(const std::vector<Atom>& atoms);

// One-electron h = T + V and S through the engine.
void one_electron(GpuHF_IBS& eng, const std::vector<Atom>& atoms, Mat& S, Mat& h);

// Superposition of atomic densities (the SAD seed): each atom's spherically averaged, fractional-
// occupation neutral-atom density in the atom's own shells, placed block-diagonally.  This is the
// matrix form of the seed; the reference's SAD (ChargeDensity/AtomicDensity.C) reads tabulated
// radial densities instead.  Atoms are matched to shells by centre.
Mat sad_density(const Basis& b, const std::vector<Atom>& atoms, const qcjk_opts& o);
// The same for many molecules at once; the atomic SCFs not yet cached run as one batch.
std::vector<Mat> sad_densities(const std::vector<const Basis*>& bases, const std::vector<std::vector<Atom>>& atoms,
                               const qcjk_opts& o);

// RHF to convergence.  D0 (optional) is an explicit seed density; otherwise the core guess.
SCFResult rhf(GpuHF_IBS& eng, const std::vector<Atom>& atoms, int nelec, const SCFParams& par, const Mat* D0 = nullptr,
              const Basis* basis_for_sad = nullptr);

// Many independent molecules, each running the rhf() recipe, with ONE batched J/K build per
// iteration for all of them (qcjk_create_batch; the build replays a captured hipGraph).  A
// molecule that has converged keeps its density; the loop ends when all have converged.
// Per-molecule host work (Fock assembly, DIIS, diagonalisation) runs on nthreads threads.
struct BatchSCFResult {
    std::vector<SCFResult> mols;
    int iterations = 0;                  // batched builds
    double t_total = 0, t_setup = 0, t_jk = 0, t_host = 0;
    double jk_device_ms = 0;             // device time of the builds
    long long quartets = 0, quartets_fp32 = 0;
};
BatchSCFResult rhf_batch(const std::vector<const Basis*>& bases, const std::vector<std::vector<Atom>>& atoms,
                         const std::vector<int>& nelec, const SCFParams& par, const qcjk_opts& opts, unsigned jk_flags,
                         int nthreads, const std::vector<Mat>* seeds = nullptr);

}  // namespace qchem_amd
