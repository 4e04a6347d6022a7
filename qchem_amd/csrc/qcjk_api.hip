// qcjk_api.hip -- the C ABI declared in include/qcjk.h: handle life cycle, work lists, launches.
//
// Host-side responsibilities:
//   * validate the flat basis (one molecule, or a batch of molecules sharing one handle), regenerate
//     the Boys table (positive-term series in long double),
//   * enumerate each molecule's unique shell pairs (first shell has the higher angular momentum),
//     group them by pair class and molecule, run the pair-precompute and Schwarz kernels, sort each
//     molecule's run of a class by the bound,
//   * build, per (bra class >= ket class), the work-unit prefix sums the kernels binary-search,
//   * per build: prescale D' = ns (x) ns o D, zero the accumulators, launch the 21 class kernels
//     (one launch per class for the whole batch), symmetrise J = Jt + Jt^T (and K) with the ns
//     post-scale; optionally as one captured hipGraph replayed per build.
// No CPU fallback exists: every path needs the HIP device.
#include "qcjk.h"
#include "qcjk_internal.h"
#include "mnd_device.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

using namespace qcjk;

namespace {

constexpr double kPruneTol = 1e-30;  // create-time: drop quartets whose Schwarz product is below this
constexpr double kFp32MaxLog10 = 30.0;  // FP32 quartets only where (2 alpha)^L < 1e30 (FP32 max 3.4e38)

// Variants: 0 = lane per quartet (64 ket pairs per unit), 1 = wavefront per quartet with LDS
// transforms, 2 = wavefront per quartet with register transforms (1 ket pair per unit),
// 3 = lane per quartet chunked over the ket's first component (64 per unit), 4 = variant 2 with the
// sum-factorised bra transform (1 per unit), 5 and 6 = variants 2 and 4 with small LDS batches,
// 7 and 8 = variant 2 with both transforms sum-factorised, big and small LDS batches, 9 and 10 =
// variant 2 in lane-R form (whole R tables per lane), R batches of ~16 KB and of ~8 KB with at most
// 256 registers.
constexpr int kNumVariants = 11;
constexpr int kTile[kNumVariants] = {64, 1, 1, 64, 1, 1, 1, 1, 1, 1, 1};

struct ClassPairWork {
    int bra, ket;
    int variant = 1;
    const ClassKernelInfo* info;
    long long nunits[2] = {0, 0};
    long long nquartets = 0;   // surviving at create time
    long long* d_start[2] = {nullptr, nullptr};
    int2* d_kseg = nullptr;
    double prim_flops = 0;     // per primitive quartet
    double digest_flops = 0;   // per contracted quartet
    bool fp32_ok = true;       // (2 alpha_max)^L of the class fits FP32 with headroom
    bool has(int v) const
    {
        return v == 1 || v == 2 || v >= 4 || (v == 0 && info->light != nullptr) || (v == 3 && info->lightc != nullptr);
    }
    ClassLauncher launcher(int v) const
    {
        switch (v) {
        case 0: return info->light;
        case 1: return info->heavy;
        case 2: return info->heavy2;
        case 3: return info->lightc;
        case 4: return info->heavy2f;
        case 5: return info->heavy2s;
        case 6: return info->heavy2fs;
        case 7: return info->heavy2k;
        case 8: return info->heavy2ks;
        case 9: return info->heavy2r;
        default: return info->heavy2rw;
        }
    }
    int unit_list(int v) const { return (v == 0 || v == 3) ? 0 : 1; }
};

// Boys table F_j(0.1 k), k = 0..400, j = 0..23 (positive-term series, long double), computed once
// per process and shared by every handle.
void make_boys_table_uncached(std::vector<double>& tab);
void make_boys_table(std::vector<double>& tab)
{
    static const std::vector<double> cached = [] {
        std::vector<double> t;
        make_boys_table_uncached(t);
        return t;
    }();
    tab = cached;
}

void make_boys_table_uncached(std::vector<double>& tab)
{
    tab.assign(kBoysRows * kBoysCols, 0.0);
    for (int it = 0; it < kBoysRows; ++it) {
        const long double T = 0.1L * it;
        const long double eT = expl(-T);
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
}

template <typename T>
hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

// Per molecule (blockIdx.y): mol = {element offset, n_m, first basis function, 0}.
__global__ void k_prescale(const double* __restrict__ D, const double* __restrict__ ns, const int4* __restrict__ mol,
                           double* __restrict__ Dp, double* __restrict__ Jt, double* __restrict__ Kt,
                           unsigned long long* __restrict__ dmax_bits)
{
    const int4 M = mol[blockIdx.y];
    const int n = M.y;
    const long long nn = (long long)n * n;
    double m = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn; i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / n), c = (int)(i % n);
        const double v = D[M.x + i] * ns[M.z + r] * ns[M.z + c];
        Dp[M.x + i] = v;
        Jt[M.x + i] = 0.0;
        Kt[M.x + i] = 0.0;
        m = fmax(m, fabs(v));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(dmax_bits, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_finalize(const double* __restrict__ Jt, const double* __restrict__ Kt, const double* __restrict__ ns,
                           const int4* __restrict__ mol, double* __restrict__ J, double* __restrict__ K)
{
    const int4 M = mol[blockIdx.y];
    const int n = M.y;
    const long long nn = (long long)n * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn; i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / n), c = (int)(i % n);
        const long long t = M.x + (long long)c * n + r;
        const double w = ns[M.z + r] * ns[M.z + c];
        if (J) J[M.x + i] = (Jt[M.x + i] + Jt[t]) * w;
        if (K) K[M.x + i] = (Kt[M.x + i] + Kt[t]) * w;
    }
}

}  // namespace

constexpr int kMaxStreams = 4;  // class launches spread over the build stream and up to 3 side streams

struct qcjk_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nstreams = 1;                            // concurrent class streams (qcjk_set_streams)
    hipStream_t side[kMaxStreams - 1] = {};      // default priority, non-blocking
    hipEvent_t fork = nullptr, join[kMaxStreams - 1] = {};
    qcjk_opts opts{};
    int n = 0, nshell = 0;           // basis functions and shells over all molecules
    int nmol = 1;
    long long nelem = 0;             // packed matrix elements, sum of n_m^2
    std::vector<int> mol_shell0;     // [nmol+1] first shell of each molecule
    std::vector<int4> mol;           // {element offset, n_m, first basis function, 0}
    int max_nm = 0;
    std::vector<int> sl, snp, spoff, sbf, smol;  // sbf: offset within the shell's molecule
    std::vector<double> scen, exps, coefs, ns;
    // pairs, sorted: class c occupies [coff[c], coff[c+1])
    int coff[kNumPairClasses + 1] = {};
    std::vector<int> pair_first, pair_second;  // shell indices per sorted pair
    std::vector<double> qraw, qnorm;
    std::vector<int> pair_np;
    long long npairs = 0;
    // device
    double* d_boys = nullptr;
    double *d_exps = nullptr, *d_coefs = nullptr, *d_cen = nullptr, *d_ns = nullptr;
    int *d_spoff = nullptr, *d_snp = nullptr;
    double4 *d_pp0 = nullptr, *d_pp1 = nullptr;
    int4* d_pinfo = nullptr;
    double4* d_pgeo = nullptr;
    int2* d_pmat = nullptr;
    int4* d_mol = nullptr;
    double *d_D = nullptr, *d_Dp = nullptr, *d_Jt = nullptr, *d_Kt = nullptr, *d_J = nullptr, *d_K = nullptr;
    unsigned long long* d_dmax = nullptr;
    unsigned long long* d_counters = nullptr;  // 3 per class pair
    std::vector<ClassPairWork> work;
    // per-class profiling
    bool profiling = false;
    std::vector<hipEvent_t> pev;            // 2 per class pair
    std::vector<qcjk_class_stat> ptot;      // running totals
    bool pending = false;                   // events of the last build not yet folded into ptot
    unsigned pflags = 0;
    // captured build (qcjk_graph_build): replayed while pointers, flags and variants are unchanged
    hipGraphExec_t gexec = nullptr;
    const double* gD = nullptr;
    double *gJ = nullptr, *gK = nullptr;
    unsigned gflags = 0;
};

#define HIPCHK(x)                                          \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return QCJK_ERR_HIP;         \
    } while (0)

static void free_handle(qcjk_handle* h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    for (auto& w : h->work) { hipFree(w.d_start[0]); hipFree(w.d_start[1]); hipFree(w.d_kseg); }
    void* ptrs[] = {h->d_boys, h->d_exps, h->d_coefs, h->d_cen, h->d_ns, h->d_spoff, h->d_snp, h->d_pp0, h->d_pp1,
                    h->d_pinfo, h->d_pgeo, h->d_pmat, h->d_mol, h->d_D, h->d_Dp, h->d_Jt, h->d_Kt, h->d_J, h->d_K,
                    h->d_dmax, h->d_counters};
    for (void* p : ptrs) if (p) hipFree(p);
    for (hipEvent_t e : h->pev) hipEventDestroy(e);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->fork) hipEventDestroy(h->fork);
    for (int k = 0; k < kMaxStreams - 1; ++k) {
        if (h->join[k]) hipEventDestroy(h->join[k]);
        if (h->side[k]) hipStreamDestroy(h->side[k]);
    }
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

static BuildArgs make_args(qcjk_handle* h, unsigned flags)
{
    BuildArgs a{};
    a.D = h->d_Dp;
    a.Jt = h->d_Jt;
    a.Kt = h->d_Kt;
    a.pmat = h->d_pmat;
    a.flags = flags;
    a.boys = h->d_boys;
    a.pp0 = h->d_pp0;
    a.pp1 = h->d_pp1;
    a.pinfo = h->d_pinfo;
    a.pgeo = h->d_pgeo;
    a.dmax = reinterpret_cast<const double*>(h->d_dmax);
    a.screen_tol = h->opts.screen_tol;
    a.mixed_tol = h->opts.mixed_tol;
    a.counters = nullptr;
    return a;
}

static int create_impl(const qcjk_basis* mols, int nmol, const qcjk_opts* o, qcjk_handle* h)
{
    h->opts = *o;
    h->device = o->device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return QCJK_ERR_NODEVICE;
    if (o->device < 0 || o->device >= ndev) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(o->device));
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    HIPCHK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
    for (int k = 0; k < kMaxStreams - 1; ++k) {
        HIPCHK(hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&h->join[k], hipEventDisableTiming));
    }
    h->nstreams = kMaxStreams - 1;
    const bool force_heavy = std::getenv("QCJK_FORCE_HEAVY") != nullptr;

    // concatenate the molecules' shells; basis functions are numbered within each molecule
    h->nmol = nmol;
    h->mol_shell0.assign(nmol + 1, 0);
    h->mol.resize(nmol);
    int np = 0, nbf = 0, ns = 0;
    long long nelem = 0;
    for (int m = 0; m < nmol; ++m) {
        const qcjk_basis* b = &mols[m];
        h->mol_shell0[m] = ns;
        int nm = 0, npm = 0;
        for (int s = 0; s < b->nshell; ++s) {
            h->sl.push_back(b->shell_l[s]);
            h->snp.push_back(b->shell_nprim[s]);
            h->spoff.push_back(np + npm);
            h->sbf.push_back(nm);
            h->smol.push_back(m);
            for (int x = 0; x < 3; ++x) h->scen.push_back(b->shell_center[3 * s + x]);
            npm += b->shell_nprim[s];
            nm += ncart(b->shell_l[s]);
        }
        // exponents, coefficients and ns straight from the caller's arrays (shell order)
        h->exps.insert(h->exps.end(), b->prim_exp, b->prim_exp + npm);
        h->coefs.insert(h->coefs.end(), b->prim_coef, b->prim_coef + npm);
        h->ns.insert(h->ns.end(), b->bf_norm, b->bf_norm + nm);
        np += npm;
        h->mol[m] = make_int4((int)nelem, nm, nbf, 0);
        h->max_nm = std::max(h->max_nm, nm);
        nelem += (long long)nm * nm;
        nbf += nm;
        ns += b->nshell;
    }
    h->mol_shell0[nmol] = ns;
    if (nelem > 0x7fffffffLL) return QCJK_ERR_UNSUPPORTED;  // element offsets are 32-bit
    h->nshell = ns;
    h->n = nbf;
    h->nelem = nelem;
    for (int i = 0; i < np; ++i)
        if (!(h->exps[i] > 0.0) || !std::isfinite(h->coefs[i])) return QCJK_ERR_ARG;
    for (int i = 0; i < nbf; ++i)
        if (!(h->ns[i] > 0.0) || !std::isfinite(h->ns[i])) return QCJK_ERR_ARG;

    // Boys table + basis arrays
    std::vector<double> tab;
    make_boys_table(tab);
    HIPCHK(dalloc(&h->d_boys, tab.size()));
    HIPCHK(hipMemcpy(h->d_boys, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(dalloc(&h->d_exps, np));
    HIPCHK(dalloc(&h->d_coefs, np));
    HIPCHK(dalloc(&h->d_cen, 3 * ns));
    HIPCHK(dalloc(&h->d_ns, nbf));
    HIPCHK(dalloc(&h->d_spoff, ns));
    HIPCHK(dalloc(&h->d_snp, ns));
    HIPCHK(dalloc(&h->d_mol, nmol));
    HIPCHK(hipMemcpy(h->d_exps, h->exps.data(), np * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_coefs, h->coefs.data(), np * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_cen, h->scen.data(), 3 * ns * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_ns, h->ns.data(), nbf * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_spoff, h->spoff.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_snp, h->snp.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_mol, h->mol.data(), nmol * sizeof(int4), hipMemcpyHostToDevice));

    // unique shell pairs of each molecule, oriented (higher l first; equal l: higher index first),
    // grouped by class, molecule order within a class
    std::vector<std::vector<int>> cls_first(kNumPairClasses), cls_second(kNumPairClasses);
    for (int m = 0; m < nmol; ++m)
        for (int A = h->mol_shell0[m]; A < h->mol_shell0[m + 1]; ++A)
            for (int B = h->mol_shell0[m]; B <= A; ++B) {
                int f = A, s = B;
                if (h->sl[A] < h->sl[B]) std::swap(f, s);
                const int c = pair_class(h->sl[f], h->sl[s]);
                cls_first[c].push_back(f);
                cls_second[c].push_back(s);
            }
    std::vector<int> pf, psd;
    for (int c = 0; c < kNumPairClasses; ++c) {
        h->coff[c] = (int)pf.size();
        pf.insert(pf.end(), cls_first[c].begin(), cls_first[c].end());
        psd.insert(psd.end(), cls_second[c].begin(), cls_second[c].end());
    }
    h->coff[kNumPairClasses] = (int)pf.size();
    const int npair = (int)pf.size();
    h->npairs = npair;

    // primitive-pair precompute
    std::vector<int4> pshell(npair);
    std::vector<int2> pnp(npair);
    std::vector<int> ppoff(npair);
    long long npp = 0;
    for (int k = 0; k < npair; ++k) {
        pshell[k] = make_int4(pf[k], psd[k], h->spoff[pf[k]], h->spoff[psd[k]]);
        pnp[k] = make_int2(h->snp[pf[k]], h->snp[psd[k]]);
        ppoff[k] = (int)npp;
        npp += (long long)h->snp[pf[k]] * h->snp[psd[k]];
    }
    if (npp > 0x7fffffffLL) return QCJK_ERR_UNSUPPORTED;
    int4* d_pshell = nullptr;
    int2* d_pnp = nullptr;
    int* d_ppoff = nullptr;
    HIPCHK(dalloc(&d_pshell, npair));
    HIPCHK(dalloc(&d_pnp, npair));
    HIPCHK(dalloc(&d_ppoff, npair));
    HIPCHK(dalloc(&h->d_pp0, npp));
    HIPCHK(dalloc(&h->d_pp1, npp));
    HIPCHK(hipMemcpy(d_pshell, pshell.data(), npair * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_pnp, pnp.data(), npair * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ppoff, ppoff.data(), npair * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(launch_pair_setup(d_pshell, d_pnp, h->d_exps, h->d_coefs, h->d_cen, npair, h->d_pp0, h->d_pp1, d_ppoff,
                             h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    hipFree(d_pshell);
    hipFree(d_pnp);
    hipFree(d_ppoff);

    // pair records.  The Schwarz kernels index ns by the basis-function offsets, so they get the
    // global offsets; the J/K kernels get offsets within the molecule plus the molecule's block.
    auto make_pinfo = [&](int k, bool global_bf) {
        const int f = pf[k], s = psd[k];
        const int npr = h->snp[f] * h->snp[s];
        const int base = global_bf ? h->mol[h->smol[f]].z : 0;
        return make_int4(base + h->sbf[f], base + h->sbf[s], ppoff[k], npr | ((f == s ? 1 : 0) << 16));
    };
    auto make_pgeo = [&](int k, double q) {
        const int f = pf[k], s = psd[k];
        return make_double4(h->scen[3 * f] - h->scen[3 * s], h->scen[3 * f + 1] - h->scen[3 * s + 1],
                            h->scen[3 * f + 2] - h->scen[3 * s + 2], q);
    };
    std::vector<int4> pinfo(npair);
    std::vector<double4> pgeo(npair);
    for (int k = 0; k < npair; ++k) {
        if (h->snp[pf[k]] * h->snp[psd[k]] > 0xffff) return QCJK_ERR_UNSUPPORTED;
        pinfo[k] = make_pinfo(k, true);
        pgeo[k] = make_pgeo(k, 0.0);
    }
    HIPCHK(dalloc(&h->d_pinfo, npair));
    HIPCHK(dalloc(&h->d_pgeo, npair));
    HIPCHK(dalloc(&h->d_pmat, npair));
    HIPCHK(hipMemcpy(h->d_pinfo, pinfo.data(), npair * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_pgeo, pgeo.data(), npair * sizeof(double4), hipMemcpyHostToDevice));
    double *d_qraw = nullptr, *d_qnorm = nullptr;
    HIPCHK(dalloc(&d_qraw, npair));
    HIPCHK(dalloc(&d_qnorm, npair));
    {
        BuildArgs a = make_args(h, 0);
        for (int c = 0; c < kNumPairClasses; ++c) {
            const int off = h->coff[c], cnt = h->coff[c + 1] - h->coff[c];
            HIPCHK(launch_schwarz(c, a, off, cnt, h->d_ns, d_qraw + off, d_qnorm + off, h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> qraw(npair), qnorm(npair);
    HIPCHK(hipMemcpy(qraw.data(), d_qraw, npair * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(qnorm.data(), d_qnorm, npair * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(d_qraw);
    hipFree(d_qnorm);

    // sort each class by molecule, then descending raw bound; rebuild records in sorted order
    std::vector<int> order(npair);
    std::iota(order.begin(), order.end(), 0);
    for (int c = 0; c < kNumPairClasses; ++c)
        std::stable_sort(order.begin() + h->coff[c], order.begin() + h->coff[c + 1], [&](int x, int y) {
            const int mx = h->smol[pf[x]], my = h->smol[pf[y]];
            return mx != my ? mx < my : qraw[x] > qraw[y];
        });
    h->pair_first.resize(npair);
    h->pair_second.resize(npair);
    h->qraw.resize(npair);
    h->qnorm.resize(npair);
    h->pair_np.resize(npair);
    std::vector<int2> pmat(npair);
    std::vector<int> pmol(npair);
    for (int k = 0; k < npair; ++k) {
        const int src = order[k];
        pinfo[k] = make_pinfo(src, false);
        pgeo[k] = make_pgeo(src, qraw[src]);
        h->pair_first[k] = pf[src];
        h->pair_second[k] = psd[src];
        h->qraw[k] = qraw[src];
        h->qnorm[k] = qnorm[src];
        h->pair_np[k] = h->snp[pf[src]] * h->snp[psd[src]];
        pmol[k] = h->smol[pf[src]];
        pmat[k] = make_int2(h->mol[pmol[k]].x, h->mol[pmol[k]].y);
    }
    HIPCHK(hipMemcpy(h->d_pinfo, pinfo.data(), npair * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_pgeo, pgeo.data(), npair * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_pmat, pmat.data(), npair * sizeof(int2), hipMemcpyHostToDevice));

    // per class: each molecule's run [mstart[c][m], mstart[c][m+1]) (class-local indices)
    std::vector<std::vector<int>> mstart(kNumPairClasses, std::vector<int>(nmol + 1, 0));
    for (int c = 0; c < kNumPairClasses; ++c) {
        std::vector<int>& st = mstart[c];
        for (int k = h->coff[c]; k < h->coff[c + 1]; ++k) ++st[pmol[k] + 1];
        for (int m = 0; m < nmol; ++m) st[m + 1] += st[m];
    }

    // largest primitive-pair exponent p = a_i + b_j per pair class (FP32 range guard)
    double class_pmax[kNumPairClasses] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < npair; ++k) {
        double amax[2] = {0.0, 0.0};
        const int sh[2] = {h->pair_first[k], h->pair_second[k]};
        for (int x = 0; x < 2; ++x)
            for (int i = 0; i < h->snp[sh[x]]; ++i) amax[x] = std::max(amax[x], h->exps[h->spoff[sh[x]] + i]);
        int c = 0;
        while (k >= h->coff[c + 1]) ++c;
        class_pmax[c] = std::max(class_pmax[c], amax[0] + amax[1]);
    }

    // work lists per (bra class >= ket class): each bra pair meets the ket pairs of its own molecule
    for (int bc = 0; bc < kNumPairClasses; ++bc)
        for (int kc = 0; kc <= bc; ++kc) {
            const int nb = h->coff[bc + 1] - h->coff[bc], nk = h->coff[kc + 1] - h->coff[kc];
            if (nb == 0 || nk == 0) continue;
            ClassPairWork w;
            w.bra = bc;
            w.ket = kc;
            w.info = class_kernel(bc, kc);
            if (!w.info) return QCJK_ERR_UNSUPPORTED;
            w.variant = force_heavy ? 1 : ((w.info->tile == 64 && w.info->light) ? 0 : 10);
            {
                // FP32 range guard: R^{(0)}_{tuv} carries up to (2 alpha)^L with alpha <= min(p, q).
                const int L = w.info->la + w.info->lb + w.info->lc + w.info->ld;
                const double amax = std::min(class_pmax[bc], class_pmax[kc]);
                w.fp32_ok = L * std::log10(std::max(2.0 * amax, 1.0)) < kFp32MaxLog10;
            }
            w.prim_flops = prim_flops(w.info->la, w.info->lb, w.info->lc, w.info->ld);
            w.digest_flops = digest_flops(w.info->la, w.info->lb, w.info->lc, w.info->ld);
            std::vector<long long> start0(nb + 1, 0), start1(nb + 1, 0);
            std::vector<int2> kseg(nb);
            const double* qk = h->qraw.data() + h->coff[kc];
            for (int P = 0; P < nb; ++P) {
                const int m = pmol[h->coff[bc] + P];
                const int k0 = mstart[kc][m], k1 = mstart[kc][m + 1];
                const double qp = h->qraw[h->coff[bc] + P];
                // the molecule's ket run is sorted descending: survivors are a prefix
                int lo = k0, hi = k1;
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (qp * qk[mid] >= kPruneTol) lo = mid + 1;
                    else hi = mid;
                }
                int cnt = lo - k0;
                if (bc == kc) cnt = std::min(cnt, P - k0 + 1);  // ket index <= bra index
                kseg[P] = make_int2(k0, std::max(cnt, 0));
                start0[P + 1] = start0[P] + (kseg[P].y + kTile[0] - 1) / kTile[0];
                start1[P + 1] = start1[P] + kseg[P].y;
                w.nquartets += kseg[P].y;
            }
            w.nunits[0] = start0[nb];
            w.nunits[1] = start1[nb];
            HIPCHK(dalloc(&w.d_start[0], nb + 1));
            HIPCHK(dalloc(&w.d_start[1], nb + 1));
            HIPCHK(dalloc(&w.d_kseg, nb));
            HIPCHK(hipMemcpy(w.d_start[0], start0.data(), (nb + 1) * sizeof(long long), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(w.d_start[1], start1.data(), (nb + 1) * sizeof(long long), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(w.d_kseg, kseg.data(), nb * sizeof(int2), hipMemcpyHostToDevice));
            h->work.push_back(w);
        }

    const size_t nn = (size_t)nelem;
    HIPCHK(dalloc(&h->d_D, nn));
    HIPCHK(dalloc(&h->d_Dp, nn));
    HIPCHK(dalloc(&h->d_Jt, nn));
    HIPCHK(dalloc(&h->d_Kt, nn));
    HIPCHK(dalloc(&h->d_J, nn));
    HIPCHK(dalloc(&h->d_K, nn));
    HIPCHK(dalloc(&h->d_dmax, 1));
    HIPCHK(dalloc(&h->d_counters, 3 * h->work.size() + 3));
    return QCJK_OK;
}

static ClassPairArgs class_args(const qcjk_handle* h, const ClassPairWork& w, int variant)
{
    ClassPairArgs c{};
    c.bra_off = h->coff[w.bra];
    c.nbra = h->coff[w.bra + 1] - h->coff[w.bra];
    c.ket_off = h->coff[w.ket];
    c.nket = h->coff[w.ket + 1] - h->coff[w.ket];
    c.same = w.bra == w.ket;
    c.unit_start = w.d_start[w.unit_list(variant)];
    c.kseg = w.d_kseg;
    c.nunits = w.nunits[w.unit_list(variant)];
    c.tile = kTile[variant];
    return c;
}

// Prescale D, zero the accumulators and counters (start of every build).
// Elementwise kernels over the packed matrices: grid (x blocks per molecule, molecules).
static dim3 elem_grid(const qcjk_handle* h)
{
    const long long per = ((long long)h->max_nm * h->max_nm + 255) / 256;
    const long long cap = std::max<long long>(1, 4096 / h->nmol);
    return dim3((unsigned)std::max<long long>(1, std::min(per, cap)), (unsigned)h->nmol);
}

// Prescale D, zero the accumulators and counters (start of every build).  Kernels only, so a
// build can be captured into a hipGraph.
static int begin_build(qcjk_handle* h, const double* dD, hipStream_t s)
{
    const int nctr = (int)(3 * h->work.size() + 3);
    hipLaunchKernelGGL(k_zero_u64, dim3((nctr + 255) / 256), dim3(256), 0, s, h->d_counters, nctr);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, h->d_dmax, 1);
    hipLaunchKernelGGL(k_prescale, elem_grid(h), dim3(256), 0, s, dD, h->d_ns, h->d_mol, h->d_Dp, h->d_Jt, h->d_Kt,
                       h->d_dmax);
    HIPCHK(hipGetLastError());
    return QCJK_OK;
}

// Fold the previous profiled build's events and counters into the running totals.
static int fold_profile(qcjk_handle* h)
{
    if (!h->pending) return QCJK_OK;
    h->pending = false;
    const int nctr = (int)(3 * h->work.size() + 3);
    std::vector<unsigned long long> ctr(nctr);
    for (size_t i = 0; i < h->work.size(); ++i) HIPCHK(hipEventSynchronize(h->pev[2 * i + 1]));  // every stream
    HIPCHK(hipMemcpy(ctr.data(), h->d_counters, nctr * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h->work.size(); ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->pev[2 * i], h->pev[2 * i + 1]));
        qcjk_class_stat& t = h->ptot[i];
        t.launches += 1;
        t.quartets += (long long)ctr[3 * i];
        t.prim_quartets += (long long)ctr[3 * i + 2];
        t.flops += (double)ctr[3 * i + 2] * h->work[i].prim_flops + (double)ctr[3 * i] * h->work[i].digest_flops;
        t.device_ms += ms;
    }
    return QCJK_OK;
}

extern "C" {

int qcjk_profile(qcjk_handle* h, int enable)
{
    if (!h) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(h->device));
    if (h->pev.empty()) {
        h->pev.resize(2 * h->work.size());
        for (auto& e : h->pev) HIPCHK(hipEventCreate(&e));
    }
    h->pending = false;
    h->profiling = enable != 0;
    h->ptot.assign(h->work.size(), qcjk_class_stat{});
    for (size_t i = 0; i < h->work.size(); ++i) {
        const ClassKernelInfo* k = h->work[i].info;
        h->ptot[i].la = k->la; h->ptot[i].lb = k->lb; h->ptot[i].lc = k->lc; h->ptot[i].ld = k->ld;
        h->ptot[i].wavefront_per_quartet = h->work[i].variant;
        h->ptot[i].alt_ms = -1.0;
    }
    return QCJK_OK;
}

int qcjk_class_report(qcjk_handle* h, qcjk_class_stat* out, int max, int* count)
{
    if (!h || !count) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(h->device));
    const int rc = fold_profile(h);
    if (rc != QCJK_OK) return rc;
    *count = (int)h->ptot.size();
    for (int i = 0; i < (int)h->ptot.size() && i < max && out; ++i) out[i] = h->ptot[i];
    return QCJK_OK;
}

static int check_basis(const qcjk_basis* basis)
{
    if (!basis || basis->nshell <= 0 || !basis->shell_l || !basis->shell_nprim || !basis->shell_center ||
        !basis->prim_exp || !basis->prim_coef || !basis->bf_norm)
        return QCJK_ERR_ARG;
    if (basis->order != 0) return QCJK_ERR_UNSUPPORTED;
    for (int s = 0; s < basis->nshell; ++s) {
        if (basis->shell_l[s] < 0 || basis->shell_l[s] > QCJK_MAX_L) return QCJK_ERR_UNSUPPORTED;
        if (basis->shell_nprim[s] < 1 || basis->shell_nprim[s] > QCJK_MAX_NPRIM) return QCJK_ERR_UNSUPPORTED;
    }
    return QCJK_OK;
}

int qcjk_create_batch(const qcjk_basis* mols, int nmol, const qcjk_opts* opts, qcjk_handle** out)
{
    if (!out) return QCJK_ERR_ARG;
    *out = nullptr;
    if (!mols || nmol < 1 || nmol > 65535) return QCJK_ERR_ARG;
    for (int m = 0; m < nmol; ++m) {
        const int rc = check_basis(&mols[m]);
        if (rc != QCJK_OK) return rc;
    }
    qcjk_opts o{};
    if (opts) o = *opts;
    qcjk_handle* h = new (std::nothrow) qcjk_handle();
    if (!h) return QCJK_ERR_NOMEM;
    const int rc = create_impl(mols, nmol, &o, h);
    if (rc != QCJK_OK) {
        free_handle(h);
        return rc;
    }
    *out = h;
    return QCJK_OK;
}

int qcjk_create(const qcjk_basis* basis, const qcjk_opts* opts, qcjk_handle** out)
{
    return qcjk_create_batch(basis, 1, opts, out);
}

int qcjk_nbf(const qcjk_handle* h) { return h ? h->n : -1; }

int qcjk_layout(const qcjk_handle* h, int32_t* nmol, int32_t* nbf, int64_t* offset, int64_t* nelem)
{
    if (!h) return QCJK_ERR_ARG;
    if (nmol) *nmol = h->nmol;
    if (nelem) *nelem = h->nelem;
    for (int m = 0; m < h->nmol; ++m) {
        if (nbf) nbf[m] = h->mol[m].y;
        if (offset) offset[m] = h->mol[m].x;
    }
    return QCJK_OK;
}

// Enqueue one build on s: prescale, the class launches, symmetrise.  Kernels only.
static int enqueue_build(qcjk_handle* h, const double* dD, double* dJ, double* dK, unsigned flags, hipStream_t s)
{
    {
        const int rc = begin_build(h, dD, s);
        if (rc != QCJK_OK) return rc;
    }
    if (flags & (QCJK_J | QCJK_K)) {
        // Class launches round-robin over ns streams forked from and joined back into s, so one
        // class's tail overlaps the next class's start (every kernel adds into Jt/Kt atomically).
        // The classes are taken alternately from the light and the heavy end of the class list,
        // so each stream mixes latency-bound lane kernels with issue-bound wavefront kernels.
        // Per-class profiling events go on the class's own stream.
        const int ns = h->nstreams;
        const size_t nw = h->work.size();
        if (ns > 1) {
            HIPCHK(hipEventRecord(h->fork, s));
            for (int k = 0; k < ns - 1; ++k) HIPCHK(hipStreamWaitEvent(h->side[k], h->fork, 0));
        }
        BuildArgs a = make_args(h, flags);
        for (size_t j = 0; j < nw; ++j) {
            const size_t i = (ns > 1) ? ((j & 1) ? nw - 1 - j / 2 : j / 2) : j;
            const ClassPairWork& w = h->work[i];
            const ClassPairArgs c = class_args(h, w, w.variant);
            const int k = (int)(j % (size_t)ns);
            hipStream_t si = k == 0 ? s : h->side[k - 1];
            a.counters = h->d_counters + 3 * i;
            a.flags = w.fp32_ok ? flags : (flags & ~(unsigned)QCJK_MIXED);
            if (h->profiling) HIPCHK(hipEventRecord(h->pev[2 * i], si));
            HIPCHK(w.launcher(w.variant)(a, c, si));
            if (h->profiling) HIPCHK(hipEventRecord(h->pev[2 * i + 1], si));
        }
        for (int k = 0; k < ns - 1; ++k) {
            HIPCHK(hipEventRecord(h->join[k], h->side[k]));
            HIPCHK(hipStreamWaitEvent(s, h->join[k], 0));
        }
    }
    hipLaunchKernelGGL(k_finalize, elem_grid(h), dim3(256), 0, s, h->d_Jt, h->d_Kt, h->d_ns, h->d_mol,
                       (flags & QCJK_J) ? dJ : nullptr, (flags & QCJK_K) ? dK : nullptr);
    HIPCHK(hipGetLastError());
    return QCJK_OK;
}

static int read_stats(qcjk_handle* h, qcjk_stats* stats)
{
    const int nctr = (int)(3 * h->work.size() + 3);
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    std::vector<unsigned long long> ctr(nctr);
    HIPCHK(hipMemcpy(ctr.data(), h->d_counters, nctr * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof(*stats));
    for (size_t i = 0; i < h->work.size(); ++i) {
        const ClassPairWork& w = h->work[i];
        stats->quartets_computed += (long long)ctr[3 * i];
        stats->quartets_fp32 += (long long)ctr[3 * i + 1];
        stats->prim_quartets += (long long)ctr[3 * i + 2];
        stats->flops += (double)ctr[3 * i + 2] * w.prim_flops + (double)ctr[3 * i] * w.digest_flops;
    }
    // unique quartets of every molecule: sum over molecules of npairs_m (npairs_m + 1) / 2
    std::vector<long long> np(h->nmol, 0);
    for (long long k = 0; k < h->npairs; ++k) ++np[h->smol[h->pair_first[k]]];
    for (int m = 0; m < h->nmol; ++m) stats->quartets_total += np[m] * (np[m] + 1) / 2;
    stats->device_ms = ms;
    return QCJK_OK;
}

int qcjk_build_device(qcjk_handle* h, const double* dD, double* dJ, double* dK, unsigned flags, void* stream,
                      qcjk_stats* stats)
{
    if (!h || !dD) return QCJK_ERR_ARG;
    if (!dJ) flags &= ~(unsigned)QCJK_J;
    if (!dK) flags &= ~(unsigned)QCJK_K;
    HIPCHK(hipSetDevice(h->device));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    if (h->profiling) {
        const int rc = fold_profile(h);
        if (rc != QCJK_OK) return rc;
    }
    if (stats) HIPCHK(hipEventRecord(h->ev0, s));
    {
        const int rc = enqueue_build(h, dD, dJ, dK, flags, s);
        if (rc != QCJK_OK) return rc;
    }
    if (h->profiling && (flags & (QCJK_J | QCJK_K))) h->pending = true;
    if (stats) {
        HIPCHK(hipEventRecord(h->ev1, s));
        HIPCHK(hipEventSynchronize(h->ev1));
        return read_stats(h, stats);
    }
    return QCJK_OK;
}

int qcjk_graph_build(qcjk_handle* h, const double* dD, double* dJ, double* dK, unsigned flags, void* stream)
{
    if (!h || !dD || h->profiling) return QCJK_ERR_ARG;
    if (!dJ) flags &= ~(unsigned)QCJK_J;
    if (!dK) flags &= ~(unsigned)QCJK_K;
    HIPCHK(hipSetDevice(h->device));
    if (!h->gexec || h->gD != dD || h->gJ != dJ || h->gK != dK || h->gflags != flags) {
        if (h->gexec) {
            HIPCHK(hipStreamSynchronize(h->stream));
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_build(h, dD, dJ, dK, flags, h->stream);
        const hipError_t ec = hipStreamEndCapture(h->stream, &g);
        if (rc != QCJK_OK) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(ec);
        const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPCHK(ei);
        h->gD = dD;
        h->gJ = dJ;
        h->gK = dK;
        h->gflags = flags;
    }
    HIPCHK(hipGraphLaunch(h->gexec, stream ? (hipStream_t)stream : h->stream));
    return QCJK_OK;
}

int qcjk_buffers(qcjk_handle* h, double** dD, double** dJ, double** dK)
{
    if (!h) return QCJK_ERR_ARG;
    if (dD) *dD = h->d_D;
    if (dJ) *dJ = h->d_J;
    if (dK) *dK = h->d_K;
    return QCJK_OK;
}

int qcjk_build(qcjk_handle* h, const double* D, double* J, double* K, unsigned flags, qcjk_stats* stats)
{
    if (!h || !D) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(h->device));
    const size_t bytes = (size_t)h->nelem * sizeof(double);
    HIPCHK(hipMemcpyAsync(h->d_D, D, bytes, hipMemcpyHostToDevice, h->stream));
    const bool wantJ = J && (flags & QCJK_J), wantK = K && (flags & QCJK_K);
    const bool graph = (flags & QCJK_GRAPH) != 0;
    flags &= ~(unsigned)QCJK_GRAPH;
    if (graph) {
        if (stats) HIPCHK(hipEventRecord(h->ev0, h->stream));
        const int rc = qcjk_graph_build(h, h->d_D, wantJ ? h->d_J : nullptr, wantK ? h->d_K : nullptr, flags, h->stream);
        if (rc != QCJK_OK) return rc;
        if (stats) {
            HIPCHK(hipEventRecord(h->ev1, h->stream));
            HIPCHK(hipEventSynchronize(h->ev1));
            const int rs = read_stats(h, stats);
            if (rs != QCJK_OK) return rs;
        }
    } else {
        const int rc = qcjk_build_device(h, h->d_D, wantJ ? h->d_J : nullptr, wantK ? h->d_K : nullptr, flags, h->stream,
                                         stats);
        if (rc != QCJK_OK) return rc;
    }
    // outputs not requested are left untouched
    if (wantJ) HIPCHK(hipMemcpyAsync(J, h->d_J, bytes, hipMemcpyDeviceToHost, h->stream));
    if (wantK) HIPCHK(hipMemcpyAsync(K, h->d_K, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return QCJK_OK;
}

int qcjk_one_electron_mol(qcjk_handle* h, int mol, int natom, const double* Z, const double* R, double* S,
                          double* T, double* V)
{
    if (!h || mol < 0 || mol >= h->nmol || natom < 0 || (natom > 0 && (!Z || !R))) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(h->device));
    const int n = h->mol[mol].y;
    const size_t nn = (size_t)n * n;
    const int s0 = h->mol_shell0[mol], s1 = h->mol_shell0[mol + 1];
    double *dZ = nullptr, *dR = nullptr, *dS = nullptr, *dT = nullptr, *dV = nullptr;
    int4* dsp = nullptr;
    int rc = QCJK_OK;
    std::vector<int4> sp;
    struct OE { int la, lb, start, count; };
    std::vector<OE> groups;
    for (int la = 0; la <= QCJK_MAX_L; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            const int start = (int)sp.size();
            for (int A = s0; A < s1; ++A)
                for (int B = s0; B < s1; ++B) {
                    if (h->sl[A] != la || h->sl[B] != lb) continue;
                    if (la == lb && B > A) continue;
                    sp.push_back(make_int4(A, B, h->sbf[A], h->sbf[B]));
                }
            groups.push_back({la, lb, start, (int)sp.size() - start});
        }
    auto fail = [&](hipError_t e) { if (e != hipSuccess && rc == QCJK_OK) rc = QCJK_ERR_HIP; };
    fail(dalloc(&dZ, natom));
    fail(dalloc(&dR, 3 * natom));
    fail(dalloc(&dS, nn));
    fail(dalloc(&dT, nn));
    fail(dalloc(&dV, nn));
    fail(dalloc(&dsp, sp.size()));
    if (rc == QCJK_OK) {
        if (natom > 0) {
            fail(hipMemcpy(dZ, Z, natom * sizeof(double), hipMemcpyHostToDevice));
            fail(hipMemcpy(dR, R, 3 * natom * sizeof(double), hipMemcpyHostToDevice));
        }
        fail(hipMemcpy(dsp, sp.data(), sp.size() * sizeof(int4), hipMemcpyHostToDevice));
        for (const OE& g : groups)
            if (rc == QCJK_OK)
                fail(launch_one_electron(g.la, g.lb, dsp + g.start, g.count, h->d_spoff, h->d_snp, h->d_exps,
                                         h->d_coefs, h->d_cen, h->d_ns + h->mol[mol].z, n, natom, dZ, dR, h->d_boys,
                                         dS, dT, dV, h->stream));
        fail(hipStreamSynchronize(h->stream));
        if (S) fail(hipMemcpy(S, dS, nn * sizeof(double), hipMemcpyDeviceToHost));
        if (T) fail(hipMemcpy(T, dT, nn * sizeof(double), hipMemcpyDeviceToHost));
        if (V) fail(hipMemcpy(V, dV, nn * sizeof(double), hipMemcpyDeviceToHost));
    }
    hipFree(dZ); hipFree(dR); hipFree(dS); hipFree(dT); hipFree(dV); hipFree(dsp);
    return rc;
}

int qcjk_one_electron(qcjk_handle* h, int natom, const double* Z, const double* R, double* S, double* T, double* V)
{
    if (!h || h->nmol != 1) return QCJK_ERR_ARG;
    return qcjk_one_electron_mol(h, 0, natom, Z, R, S, T, V);
}

int qcjk_schwarz(qcjk_handle* h, double* Q)
{
    if (!h || !Q) return QCJK_ERR_ARG;
    const int ns = h->nshell;
    for (int i = 0; i < ns * ns; ++i) Q[i] = 0.0;
    for (long long k = 0; k < h->npairs; ++k) {
        const int a = h->pair_first[k], b = h->pair_second[k];
        Q[a * ns + b] = Q[b * ns + a] = h->qnorm[k];
    }
    return QCJK_OK;
}

void qcjk_destroy(qcjk_handle* h) { free_handle(h); }

int qcjk_set_streams(qcjk_handle* h, int nstreams)
{
    if (!h || nstreams < 1 || nstreams > kMaxStreams) return QCJK_ERR_ARG;
    if (h->nstreams != nstreams) h->gD = nullptr;  // recapture the graph on its next use
    h->nstreams = nstreams;
    return QCJK_OK;
}

int qcjk_streams(const qcjk_handle* h) { return h ? h->nstreams : -1; }

int qcjk_set_variant(qcjk_handle* h, int la, int lb, int lc, int ld, int variant)
{
    if (!h || variant < 0 || variant >= kNumVariants) return QCJK_ERR_ARG;
    for (auto& w : h->work)
        if (w.info->la == la && w.info->lb == lb && w.info->lc == lc && w.info->ld == ld) {
            if (!w.has(variant)) return QCJK_ERR_UNSUPPORTED;
            if (w.variant != variant) h->gD = nullptr;  // recapture the graph on its next use
            w.variant = variant;
            return QCJK_OK;
        }
    return QCJK_OK;  // class absent from this basis: nothing to set
}

int qcjk_autotune(qcjk_handle* h, const double* dD, int reps, qcjk_class_stat* out, int max, int* count)
{
    if (!h || !dD || reps < 1) return QCJK_ERR_ARG;
    HIPCHK(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = begin_build(h, dD, s);
    const BuildArgs a = make_args(h, QCJK_J | QCJK_K);
    std::vector<qcjk_class_stat> res;
    for (size_t i = 0; i < h->work.size() && rc == QCJK_OK; ++i) {
        ClassPairWork& w = h->work[i];
        double best[kNumVariants];
        for (double& b : best) b = 1e300;
        for (int v = 0; v < kNumVariants && rc == QCJK_OK; ++v) {
            if (!w.has(v)) continue;
            const ClassPairArgs c = class_args(h, w, v);
            for (int r = 0; r < reps + 1 && rc == QCJK_OK; ++r) {  // first launch is a warm-up
                if (hipEventRecord(e0, s) != hipSuccess || w.launcher(v)(a, c, s) != hipSuccess ||
                    hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = QCJK_ERR_HIP;
                    break;
                }
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = QCJK_ERR_HIP;
                if (r > 0) best[v] = std::min(best[v], (double)ms);
            }
        }
        int bv = 0;
        for (int v = 1; v < kNumVariants; ++v)
            if (best[v] < best[bv]) bv = v;
        double alt = 1e300;
        for (int v = 0; v < kNumVariants; ++v)
            if (v != bv) alt = std::min(alt, best[v]);
        if (w.variant != bv) h->gD = nullptr;  // recapture the graph on its next use
        w.variant = bv;
        qcjk_class_stat st{};
        st.la = w.info->la; st.lb = w.info->lb; st.lc = w.info->lc; st.ld = w.info->ld;
        st.wavefront_per_quartet = w.variant;
        st.quartets = w.nquartets;
        st.device_ms = best[bv];
        st.alt_ms = alt < 1e299 ? alt : -1.0;
        res.push_back(st);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != QCJK_OK) return rc;
    if (count) *count = (int)res.size();
    for (int i = 0; i < (int)res.size() && i < max && out; ++i) out[i] = res[i];
    return QCJK_OK;
}

const char* qcjk_strerror(int code)
{
    switch (code) {
    case QCJK_OK: return "ok";
    case QCJK_ERR_ARG: return "invalid argument";
    case QCJK_ERR_UNSUPPORTED: return "unsupported basis or option";
    case QCJK_ERR_HIP: return "HIP runtime error";
    case QCJK_ERR_NOMEM: return "out of memory";
    case QCJK_ERR_NODEVICE: return "no usable HIP device";
    default: return "unknown error";
    }
}

}  // extern "C"
