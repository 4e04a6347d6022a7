/*
 * qcjk.h -- C ABI of the MI355X Fock-build engine (McMurchie-Davidson ERIs digested into J and K).
 *
 * This is the drop-in boundary for the molecular Gaussian path of JanReimers/qchem.  It replaces,
 * for that path, the ERI4 substrate behind the HF contraction faces:
 *   - Orbital_HF_IBS<T>::AccumulateDirect/AccumulateExchange      src/BasisSet/Orbital_HF_IBS.C:33-54
 *   - WholeSystemFock_IBS<T>::MakeAOFock(Dao, exchange)           src/BasisSet/Orbital_HF_IBS.C:75-87
 *   - the ERI4 build + contraction they reach:
 *       Molecule::Orbital_ERI4_IBS<E>::MakeDirect/MakeExchange     src/BasisSet/Molecule/IrrepBasisSet.C:138-183
 *       ERI4::MatMul / ScatterBoth                                 src/BasisSet/Internal/Imp/ERI4.C:10-42
 *       Orbital_ERI4_IBS<T>::Accumulate*                           src/BasisSet/Internal/Imp/Orbital_ERI4_IBS.C:46-84
 * Semantics follow Orbital_HF_IBS.C:37-42: J_ab = sum_cd (ab|cd) D_cd, K_ac = sum_bd (ab|cd) D_bd
 * (chemists' notation, K unscaled -- the -1/2 belongs to the caller's Vxc term, Hamiltonians.C:36-43).
 * The C++ adapter GpuHF_IBS (qchem_amd/host/hf_faces.hpp) turns the overwrite semantics here into the
 * reference's += semantics and its status codes into std::runtime_error.
 *
 * Conventions (all fixed by the reference):
 *   - Cartesian components per shell in reference order: for m=0..L, for l=0..L-m: (L-m-l, l, m)
 *     (PG_Cart/Imp/IrrepBasisSet.C:46-54); functions flattened shell by shell.
 *   - prim_coef are the FOLDED contraction coefficients: c_i * N(alpha_i; L,0,0) for a contracted
 *     radial, 1.0 for a single-exponent radial (PG_Cart_MnD/Imp/GaussianRF.C:557-577).
 *   - bf_norm[f] = 1/sqrt(S_ff), the per-component normalisation (PG_Cart_MnD/Imp/PGData.C:27-31).
 *   - Matrices are dense row-major n x n FP64; D must be symmetric.
 *
 * Threading: one handle = one HIP stream = one host thread at a time.  No call throws; every call
 * returns a status code (qcjk_strerror names it).  There is no CPU fallback: without a usable HIP
 * device qcjk_create fails with QCJK_ERR_NODEVICE.
 */
#ifndef QCJK_H
#define QCJK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qcjk_handle qcjk_handle;

/* Flat basis description (replaces PGData radials/pols/ns, PG_Cart_MnD/Imp/PGData.C:17-33). */
typedef struct qcjk_basis {
    int32_t       nshell;
    const int32_t* shell_l;      /* angular momentum per shell, 0..2 */
    const int32_t* shell_nprim;  /* primitives per shell, 1..QCJK_MAX_NPRIM */
    const double*  shell_center; /* 3*nshell, bohr */
    const double*  prim_exp;     /* sum(shell_nprim) exponents, shell order */
    const double*  prim_coef;    /* folded contraction coefficients, same layout */
    const double*  bf_norm;      /* ns[nbf] */
    int32_t        order;        /* 0 = reference component order (the only one supported) */
} qcjk_basis;

typedef struct qcjk_opts {
    int32_t device;      /* HIP device ordinal */
    double  screen_tol;  /* drop a shell quartet when Q_ab*Q_cd*max|D| < screen_tol (QCJK_SCREEN) */
    double  mixed_tol;   /* QCJK_MIXED: quartets with Q_ab*Q_cd*max|D| < mixed_tol run in FP32 */
    int32_t reserved[8];
} qcjk_opts;

typedef struct qcjk_stats {
    int64_t quartets_total;      /* unique contracted shell quartets (8-fold symmetry) */
    int64_t quartets_computed;   /* after screening */
    int64_t quartets_fp32;       /* of those, evaluated in FP32 (QCJK_MIXED) */
    int64_t prim_quartets;       /* primitive quartets behind quartets_computed */
    double  flops;               /* algorithmic FLOPs of the computed quartets: per primitive quartet
                                  * 40+2(L+1)+3C(L+4,4)+10+2nh(Lab)nh(Lcd)nc(c)nc(d)+2nc(a)nc(b)nh(Lab)nc(c)nc(d),
                                  * plus 12 per Cartesian element of each contracted quartet (digestion) */
    double  device_ms;           /* device time of the build, HIP events on the handle stream */
} qcjk_stats;

enum {
    QCJK_J = 1u,       /* build J */
    QCJK_K = 2u,       /* build K */
    QCJK_MIXED = 4u,   /* FP32 for quartets below opts.mixed_tol */
    QCJK_SCREEN = 8u,  /* Schwarz screening at opts.screen_tol */
    QCJK_GRAPH = 16u   /* qcjk_build only: replay the captured build (qcjk_graph_build) */
};

enum {
    QCJK_OK = 0,
    QCJK_ERR_ARG = 1,
    QCJK_ERR_UNSUPPORTED = 2,
    QCJK_ERR_HIP = 3,
    QCJK_ERR_NOMEM = 4,
    QCJK_ERR_NODEVICE = 5
};

#define QCJK_MAX_L 2
#define QCJK_MAX_NPRIM 16

/* Build the handle: uploads the basis, runs the shell-pair precompute (Hermite data, P, p, K_AB)
 * and the Schwarz-bound kernels, sorts pairs by class and bound.  Replaces the first-iteration
 * ERI4 materialisation (IrrepBasisSet.C:138-183) -- nothing O(N^4) is ever stored. */
int qcjk_create(const qcjk_basis* basis, const qcjk_opts* opts, qcjk_handle** out);

/* Number of basis functions n (over all molecules of a batch handle). */
int qcjk_nbf(const qcjk_handle* h);

/* Batch of independent molecules in one handle (no reference counterpart: the reference runs one
 * molecule per process).  Each molecule keeps its own basis, shell pairs and n_m x n_m matrices;
 * quartets never mix molecules.  Every build runs ONE launch per shell-quartet class over all
 * molecules.  D, J and K of a batch handle are PACKED: molecule m's row-major block starts at
 * element offset[m] (qcjk_layout) and the arrays hold nelem = sum n_m^2 elements; for one molecule
 * this is the plain n x n matrix.  Screening and the FP32 choice use max|D| over the batch. */
int qcjk_create_batch(const qcjk_basis* mols, int nmol, const qcjk_opts* opts, qcjk_handle** out);

/* Packed layout: molecule count, n_m and element offset per molecule ([nmol] arrays, nullable),
 * total elements.  */
int qcjk_layout(const qcjk_handle* h, int32_t* nmol, int32_t* nbf, int64_t* offset, int64_t* nelem);

/* J and/or K from a HOST density (n x n; packed blocks for a batch handle).  Requested outputs are OVERWRITTEN; an output that is null
 * or not requested in flags is not written.
 * Replaces MakeAOFock(Dao, exchange) (Orbital_HF_IBS.C:84) for both exchange=false/true at once. */
int qcjk_build(qcjk_handle* h, const double* D, double* J, double* K, unsigned flags, qcjk_stats* stats);

/* Same with DEVICE pointers already resident in HBM, on the given hipStream_t (null = the handle's
 * stream).  Asynchronous: returns after enqueueing unless stats is non-null (then it syncs to time). */
int qcjk_build_device(qcjk_handle* h, const double* dD, double* dJ, double* dK, unsigned flags,
                      void* stream, qcjk_stats* stats);

/* The per-iteration build captured once as a hipGraph and replayed (no reference counterpart; for
 * SCF loops, whose every iteration repeats the same launches).  Same arguments and results as
 * qcjk_build_device without stats; the graph is re-captured when the pointers, the flags or a
 * class's kernel variant change.  Not while per-class profiling is enabled (QCJK_ERR_ARG). */
int qcjk_graph_build(qcjk_handle* h, const double* dD, double* dJ, double* dK, unsigned flags, void* stream);

/* The handle's own device buffers for D, J and K (nelem doubles each), e.g. for qcjk_graph_build. */
int qcjk_buffers(qcjk_handle* h, double** dD, double** dJ, double** dK);

/* One-electron matrices on the device: S, T = 1/2 <-nabla^2>, V = sum_A -Z_A <a|1/|r-R_A||b>
 * (host pointers, n x n, overwritten).  Replaces the Overlap2C/Grad2/Nuclear kernels
 * (PG_Cart_MnD/Imp/GaussianRF.C:335-405) and Kinetic.C:57.  Any output may be null. */
int qcjk_one_electron(qcjk_handle* h, int natom, const double* Z, const double* R,
                      double* S, double* T, double* V);

/* Same for molecule mol of a batch handle (n_m x n_m outputs).  qcjk_one_electron is mol 0 of a
 * one-molecule handle and returns QCJK_ERR_ARG on a batch handle. */
int qcjk_one_electron_mol(qcjk_handle* h, int mol, int natom, const double* Z, const double* R,
                          double* S, double* T, double* V);

/* Schwarz bounds Q_AB = sqrt(max |(ab|ab)|) (normalised functions), nshell x nshell, host. */
int qcjk_schwarz(qcjk_handle* h, double* Q);

/* Per-class device timing (instrumentation; no reference counterpart).  While enabled, every build
 * brackets each class-pair launch with HIP events on the build stream and accumulates, per class,
 * the launches, elapsed device time and the algorithmic FLOPs of the quartets it computed. */
typedef struct qcjk_class_stat {
    int32_t la, lb, lc, ld;
    int32_t wavefront_per_quartet;  /* kernel variant in use: 0/3 lane per quartet, 1/2/4-10 wavefront per quartet */
    int64_t launches;
    int64_t quartets;
    int64_t prim_quartets;
    double  flops;
    double  device_ms;
    double  alt_ms;      /* qcjk_autotune: the other variant's time (-1 if none) */
} qcjk_class_stat;

int qcjk_profile(qcjk_handle* h, int enable);  /* enable != 0 also resets the totals */
int qcjk_class_report(qcjk_handle* h, qcjk_class_stat* out, int max, int* count);

/* Kernel variant per angular-momentum class (no reference counterpart): 0 = lane per contracted
 * quartet (register resident; compiled for classes with L <= 4), 1 = wavefront per quartet with
 * LDS-staged Hermite transforms, 2 = wavefront per quartet with register-resident transforms,
 * 4 = variant 2 with the bra transform factorised axis by axis over LDS Hermite tables, 5 and 6 =
 * variants 2 and 4 with small LDS batches (more wavefronts per SIMD), 7 and 8 = variant 2 with both
 * the bra and the ket transform factorised, big and small LDS batches, 9 and 10 = variant 2 with
 * the R tables of a batch of (bra, ket) primitive pairs built one per lane, batches of ~16 KB, and
 * of ~8 KB with at most 256 registers (1, 2, 4-10 exist for every class), 3 = lane per quartet chunked over the ket's first Cartesian component (fewer live
 * accumulators; L <= 4 classes with a p or d ket, and (dp|ds)).  qcjk_autotune times every variant
 * of every class present in this
 * basis on the given DEVICE density (reps timed launches each, after one warm-up), keeps the
 * fastest and reports, per class, the chosen variant in wavefront_per_quartet, its time in
 * device_ms and the best other variant's time in alt_ms.  Results are cached by the caller
 * (qchem_amd.tune) and re-applied with qcjk_set_variant. */
int qcjk_set_variant(qcjk_handle* h, int la, int lb, int lc, int ld, int variant);

/* Number of streams a build spreads its class launches over (no reference counterpart): 1 = all
 * on the build stream in class order; 2-4 = round-robin over the build stream and side streams
 * forked from it and joined back before the symmetrisation, so classes overlap (default 3).
 * Per-class profiling events are recorded on the class's own stream: under overlap a class's time
 * includes the time it shares the device with its neighbours. */
int qcjk_set_streams(qcjk_handle* h, int nstreams);
int qcjk_streams(const qcjk_handle* h);  /* the current count, -1 for a null handle */
int qcjk_autotune(qcjk_handle* h, const double* dD, int reps, qcjk_class_stat* out, int max, int* count);

void qcjk_destroy(qcjk_handle* h);

const char* qcjk_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* QCJK_H */
