/*
 * qchem_host.h -- C ABI of the C++20 host side (libqchem_host.so), for bindings and tools.
 *
 * The RHF driver here is the caller side of the engine boundary: it restates the reference's
 * SCF recipe (Calculation::Converge, src/Calculation/Imp/Calculation.C:166-230: core guess, DIIS
 * with FDMax = Z^2 0.1/32, SCFIterator::Iterate SCFIterator.C:240-409) and reaches J/K only through
 * the Orbital_HF_IBS / WholeSystemFock_IBS faces implemented by the GpuHF_IBS adapter.
 */
#ifndef QCHEM_HOST_H
#define QCHEM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A molecular basis as raw shells.  contracted[s] != 0 selects the contracted-radial convention
 * (coefficient * primitive norm of x^L); 0 the single-exponent radial (coefficient 1). */
typedef struct qchost_shells {
    int32_t nshell;
    const int32_t* l;
    const int32_t* nprim;
    const double* exps;
    const double* coefs;
    const int32_t* contracted;
    const double* centers; /* 3*nshell */
} qchost_shells;

typedef struct qcscf_params {
    int32_t max_iter;
    double e_tol;      /* |dE| convergence (Eh) */
    double fd_tol;     /* ||[F',D']|| convergence */
    int32_t nproj;     /* DIIS history */
    double fd_max;     /* DIIS engage threshold; <= 0 selects Z^2 0.1/32 */
    double sv_tol;     /* DIIS bordered-matrix svmin prune */
    uint32_t jk_flags; /* QCJK_J|QCJK_K plus QCJK_SCREEN / QCJK_MIXED */
    double screen_tol;
    double mixed_tol;
    int32_t device;
    int32_t verbose;
    int32_t guess;     /* seed when no D0 is given: 0 = core Hamiltonian, 1 = SAD (atomic densities) */
    int32_t incremental; /* > 0: J/K of the density change D_k - D_(k-1) added to the previous J/K (with
                          * QCJK_SCREEN the change's small max|dD| screens most quartets late in the
                          * SCF); a full build every `incremental` iterations.  0 = full builds. */
} qcscf_params;

typedef struct qcscf_result {
    double energy, e_nuc, e_one, e_coul, e_exch;
    int32_t iterations;
    int32_t converged;
    int32_t nbf;
    double t_total, t_jk, t_diag, t_setup;
    int64_t quartets, quartets_fp32;  /* shell quartets over all Fock builds; of those in FP32 */
} qcscf_result;

/* Folded coefficients and ns for the shells (the conventions the engine expects). */
int qchost_basis_prepare(const qchost_shells* sh, double* folded /* sum nprim */, double* ns /* nbf */, int32_t* nbf);

/* Closed-shell RHF.  D0 (nullable, nbf x nbf) seeds the SCF instead of the core guess.
 * D_out (nullable) receives the final density, e_hist/t_hist (nullable, max_iter) the energy and
 * elapsed seconds after every Fock build.  Returns a qcjk status code (or 100 + exception). */
int qcscf_rhf(const qchost_shells* sh, int32_t natom, const double* Z, const double* R, int32_t charge,
              const qcscf_params* par, const double* D0, qcscf_result* res, double* D_out, double* e_hist,
              double* t_hist);

/* Batched closed-shell RHF: nmol independent molecules, each running the qcscf_rhf recipe (seed:
 * D0 if given -- the molecules' n_m x n_m seeds back to back -- else par->guess), with ONE batched J/K build per iteration for all of them (qcjk_create_batch, replayed
 * hipGraph).  natom[m] atoms per molecule; Z and R hold all molecules' atoms back to back; charge
 * is per molecule (nullable = neutral).  res[m] receives each molecule's result (its t_* fields
 * are the shared batch timings); e_hist (nullable, nmol x max_iter) the energy histories; D_out
 * (nullable) the final densities, back to back like D0.
 * Per-molecule host work runs on nthreads threads.  No reference counterpart. */
typedef struct qcscf_batch_result {
    int32_t iterations;      /* batched builds */
    int32_t nconverged;
    double t_total, t_setup, t_jk, t_host;
    double jk_device_ms;     /* device time of the builds */
    int64_t quartets, quartets_fp32;
} qcscf_batch_result;

int qcscf_rhf_batch(const qchost_shells* mols, int32_t nmol, const int32_t* natom, const double* Z, const double* R,
                    const int32_t* charge, const qcscf_params* par, const double* D0, int32_t nthreads,
                    qcscf_result* res, double* e_hist, double* D_out, qcscf_batch_result* bres);

/* The SAD seed density (nbf x nbf, row-major) of a molecule: each atom's spherically averaged
 * fractional-occupation neutral-atom density in its own shells, placed block-diagonally.  Atomic
 * densities are cached per element and shell set for the life of the process. */
int qcscf_sad(const qchost_shells* sh, int32_t natom, const double* Z, const double* R, int32_t device, double* D_out);

/* Last error message of this thread (exception text), or "". */
const char* qchost_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QCHEM_HOST_H */
