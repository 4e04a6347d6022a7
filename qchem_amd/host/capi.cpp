// capi.cpp -- C ABI of the host side (include/qchem_host.h).
#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "qchem_host.h"
#include "scf.hpp"

using namespace qchem_amd;

namespace {
thread_local std::string g_err;

Basis make_basis(const qchost_shells* sh)
{
    Basis b;
    int off = 0;
    for (int s = 0; s < sh->nshell; ++s) {
        Shell x;
        x.L = sh->l[s];
        x.exps.assign(sh->exps + off, sh->exps + off + sh->nprim[s]);
        x.coefs.assign(sh->coefs + off, sh->coefs + off + sh->nprim[s]);
        x.contracted = sh->contracted[s] != 0;
        for (int k = 0; k < 3; ++k) x.R[k] = sh->centers[3 * s + k];
        off += sh->nprim[s];
        b.shells.push_back(x);
    }
    b.finalize();
    return b;
}
}  // namespace

extern "C" {

int qchost_basis_prepare(const qchost_shells* sh, double* folded, double* ns, int32_t* nbf)
{
    try {
        Basis b = make_basis(sh);
        if (folded) std::memcpy(folded, b.folded.data(), b.folded.size() * sizeof(double));
        if (ns) std::memcpy(ns, b.ns.data(), b.ns.size() * sizeof(double));
        if (nbf) *nbf = b.nbf;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 100;
    }
}

int qcscf_rhf(const qchost_shells* sh, int32_t natom, const double* Z, const double* R, int32_t charge,
              const qcscf_params* par, const double* D0, qcscf_result* res, double* D_out, double* e_hist,
              double* t_hist)
{
    try {
        Basis b = make_basis(sh);
        std::vector<Atom> atoms(natom);
        double zsum = 0.0;
        for (int i = 0; i < natom; ++i) {
            atoms[i].Z = Z[i];
            zsum += Z[i];
            for (int k = 0; k < 3; ++k) atoms[i].R[k] = R[3 * i + k];
        }
        qcjk_opts o{};
        o.device = par->device;
        o.screen_tol = par->screen_tol;
        o.mixed_tol = par->mixed_tol;
        const qcjk_basis view = b.view();
        GpuHF_IBS eng(view, o, par->jk_flags ? par->jk_flags : (QCJK_J | QCJK_K));
        SCFParams p;
        p.max_iter = par->max_iter;
        p.e_tol = par->e_tol;
        p.fd_tol = par->fd_tol;
        p.nproj = par->nproj;
        p.fd_max = par->fd_max;
        p.sv_tol = par->sv_tol;
        p.verbose = par->verbose != 0;
        p.guess = par->guess;
        p.incremental = par->incremental;
        Mat seed;
        if (D0) {
            seed = Mat(b.nbf);
            std::memcpy(seed.a.data(), D0, seed.a.size() * sizeof(double));
        }
        const int nelec = (int)std::lround(zsum) - charge;
        SCFResult r = rhf(eng, atoms, nelec, p, D0 ? &seed : nullptr, &b);
        res->energy = r.energy;
        res->e_nuc = r.e_nuc;
        res->e_one = r.e_one;
        res->e_coul = r.e_coul;
        res->e_exch = r.e_exch;
        res->iterations = r.iterations;
        res->converged = r.converged ? 1 : 0;
        res->nbf = b.nbf;
        res->t_total = r.t_total;
        res->t_jk = r.t_jk;
        res->t_diag = r.t_diag;
        res->t_setup = r.t_setup;
        res->quartets = r.quartets;
        res->quartets_fp32 = r.quartets_fp32;
        if (D_out && r.D.n == b.nbf) std::memcpy(D_out, r.D.a.data(), r.D.a.size() * sizeof(double));
        for (size_t i = 0; i < r.energy_history.size() && (int)i < par->max_iter; ++i) {
            if (e_hist) e_hist[i] = r.energy_history[i];
            if (t_hist) t_hist[i] = r.time_history[i];
        }
        return 0;
    } catch (const EngineError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 100;
    }
}

int qcscf_rhf_batch(const qchost_shells* mols, int32_t nmol, const int32_t* natom, const double* Z, const double* R,
                    const int32_t* charge, const qcscf_params* par, const double* D0, int32_t nthreads,
                    qcscf_result* res, double* e_hist, double* D_out, qcscf_batch_result* bres)
{
    try {
        if (!mols || nmol < 1 || !natom || !Z || !R || !par || !res) throw std::runtime_error("qcscf_rhf_batch: null argument");
        std::vector<Basis> bases;
        bases.reserve(nmol);
        std::vector<const Basis*> bp;
        std::vector<std::vector<Atom>> atoms(nmol);
        std::vector<int> nelec(nmol);
        int a0 = 0;
        for (int m = 0; m < nmol; ++m) {
            bases.push_back(make_basis(&mols[m]));
            double zsum = 0.0;
            atoms[m].resize(natom[m]);
            for (int i = 0; i < natom[m]; ++i) {
                atoms[m][i].Z = Z[a0 + i];
                zsum += Z[a0 + i];
                for (int k = 0; k < 3; ++k) atoms[m][i].R[k] = R[3 * (a0 + i) + k];
            }
            a0 += natom[m];
            nelec[m] = (int)std::lround(zsum) - (charge ? charge[m] : 0);
        }
        for (const Basis& b : bases) bp.push_back(&b);
        qcjk_opts o{};
        o.device = par->device;
        o.screen_tol = par->screen_tol;
        o.mixed_tol = par->mixed_tol;
        SCFParams p;
        p.max_iter = par->max_iter;
        p.e_tol = par->e_tol;
        p.fd_tol = par->fd_tol;
        p.nproj = par->nproj;
        p.fd_max = par->fd_max;
        p.sv_tol = par->sv_tol;
        p.verbose = par->verbose != 0;
        p.guess = par->guess;
        std::vector<Mat> seeds;
        if (D0) {
            size_t off = 0;
            for (const Basis& b : bases) {
                Mat s(b.nbf);
                std::memcpy(s.a.data(), D0 + off, s.a.size() * sizeof(double));
                off += s.a.size();
                seeds.push_back(std::move(s));
            }
        }
        const BatchSCFResult r = rhf_batch(bp, atoms, nelec, p, o, par->jk_flags ? par->jk_flags : (QCJK_J | QCJK_K),
                                           nthreads, D0 ? &seeds : nullptr);
        int nconv = 0;
        size_t doff = 0;
        for (int m = 0; m < nmol; ++m) {
            if (D_out && r.mols[m].D.n == bases[m].nbf)
                std::memcpy(D_out + doff, r.mols[m].D.a.data(), r.mols[m].D.a.size() * sizeof(double));
            doff += (size_t)bases[m].nbf * bases[m].nbf;
            const SCFResult& x = r.mols[m];
            qcscf_result& y = res[m];
            std::memset(&y, 0, sizeof(y));
            y.energy = x.energy;
            y.e_nuc = x.e_nuc;
            y.e_one = x.e_one;
            y.e_coul = x.e_coul;
            y.e_exch = x.e_exch;
            y.iterations = x.iterations;
            y.converged = x.converged ? 1 : 0;
            y.nbf = bases[m].nbf;
            y.t_total = r.t_total;
            y.t_jk = r.t_jk;
            y.t_diag = x.t_diag;
            y.t_setup = r.t_setup;
            nconv += x.converged ? 1 : 0;
            if (e_hist)
                for (size_t i = 0; i < x.energy_history.size() && (int)i < par->max_iter; ++i)
                    e_hist[(size_t)m * par->max_iter + i] = x.energy_history[i];
        }
        if (bres) {
            bres->iterations = r.iterations;
            bres->nconverged = nconv;
            bres->t_total = r.t_total;
            bres->t_setup = r.t_setup;
            bres->t_jk = r.t_jk;
            bres->t_host = r.t_host;
            bres->jk_device_ms = r.jk_device_ms;
            bres->quartets = r.quartets;
            bres->quartets_fp32 = r.quartets_fp32;
        }
        return 0;
    } catch (const EngineError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 100;
    }
}

int qcscf_sad(const qchost_shells* sh, int32_t natom, const double* Z, const double* R, int32_t device, double* D_out)
{
    try {
        if (!sh || natom < 1 || !Z || !R || !D_out) throw std::runtime_error("qcscf_sad: null argument");
        Basis b = make_basis(sh);
        std::vector<Atom> atoms(natom);
        for (int i = 0; i < natom; ++i) {
            atoms[i].Z = Z[i];
            for (int k = 0; k < 3; ++k) atoms[i].R[k] = R[3 * i + k];
        }
        qcjk_opts o{};
        o.device = device;
        const Mat D = sad_density(b, atoms, o);
        std::memcpy(D_out, D.a.data(), D.a.size() * sizeof(double));
        return 0;
    } catch (const EngineError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 100;
    }
}

const char* qchost_last_error(void) { return g_err.c_str(); }

}  // extern "C"
