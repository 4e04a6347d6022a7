// scf.cpp -- see scf.hpp.
#include "scf.hpp"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace qchem_amd {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// Orthogonaliser: S = U s U^T, V = U s^{-1/2} over the kept eigenvalues, Vinv = V^T S (V^T S V = 1).
struct Ortho {
    Mat V, VtS;
    int m = 0;  // kept dimension
    void set(const Mat& S)
    {
        std::vector<double> s;
        Mat U;
        sym_eigen(S, s, U);
        const int n = S.n;
        const double smax = s.empty() ? 0.0 : s.back();
        std::vector<int> keep;
        for (int k = 0; k < n; ++k)
            if (s[k] > 1e-10 * smax) keep.push_back(k);
        m = (int)keep.size();
        V = Mat(n);  // n x n storage, columns beyond m stay zero
        for (int c = 0; c < m; ++c) {
            const double f = 1.0 / std::sqrt(s[keep[c]]);
            for (int i = 0; i < n; ++i) V(i, c) = U(i, keep[c]) * f;
        }
        VtS = transpose(V) * S;
    }
    // F' = V^T F V restricted to m x m
    Mat to_ortho(const Mat& F) const
    {
        Mat full = transpose(V) * F * V;
        Mat r(m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) r(i, j) = full(i, j);
        return r;
    }
    // D' = (V^T S) D (V^T S)^T restricted to m x m
    Mat density_to_ortho(const Mat& D) const
    {
        Mat full = VtS * D * transpose(VtS);
        Mat r(m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) r(i, j) = full(i, j);
        return r;
    }
};

// Pulay DIIS on the orthonormal Fock (SCFAcceleratorDIIS.C:35-212, Math/DIIS.C:24-51).
struct DIIS {
    int nproj;
    double fd_max, sv_tol;
    std::deque<Mat> F, E;

    Mat next(const Mat& Fp, const Mat& Ep, double en)
    {
        if (en > fd_max) return Fp;  // "FD>FDMax": plain diagonalisation
        F.push_back(Fp);
        E.push_back(Ep);
        if ((int)F.size() > nproj) { F.pop_front(); E.pop_front(); }
        if (F.size() < 2) return Fp;  // "Nproj<2"
        std::vector<double> c;
        for (;;) {
            const int N = (int)F.size();
            const int nb = N + 1;
            Mat Bb(nb);
            for (int i = 0; i < N; ++i) {
                for (int j = 0; j < N; ++j) Bb(i, j) = dot(E[i], E[j]);
                Bb(i, N) = Bb(N, i) = 1.0;
            }
            // smallest singular value of the symmetric bordered matrix = min |eigenvalue|
            std::vector<double> w;
            Mat U;
            sym_eigen(Bb, w, U);
            double sv = 1e300;
            for (double x : w) sv = std::min(sv, std::fabs(x));
            if (sv < sv_tol && N >= 2) {
                F.pop_front();
                E.pop_front();
                continue;
            }
            if (nb <= 2) return Fp;  // "B.rows()<=2"
            std::vector<double> rhs(nb, 0.0), x;
            rhs[N] = 1.0;
            if (!solve(Bb.a, rhs, nb, x)) return Fp;
            c.assign(x.begin(), x.begin() + N);
            break;
        }
        Mat r(Fp.n);
        for (size_t i = 0; i < F.size(); ++i)
            for (size_t k = 0; k < r.a.size(); ++k) r.a[k] += c[i] * F[i].a[k];
        return r;
    }
};

// Linear density mixing with the adaptive relaxation of DensityMixer.C:104-139, as the
// fixed-point loop drives it (SCFIterator.C:265-351; StartingRelaxRo = 1, SCFParams.C:32):
// D_in <- r D_out + (1 - r) D_in with r = relax, or r = relax/4 and relax *= 0.8 when ||[F,D]||
// rose since the previous iteration; relax *= 1.5 when it fell; relax <= RelMax, which drops from 1
// to 0.5 once ||D_out - D_in||_F / N < 1e-5.
struct LinearMixer {
    double relax = 1.0, relmax = 1.0;
    Mat step(const Mat& Din, const Mat& Dout, double fd, double fd_old, double nelec)
    {
        if (frob(axpby(1.0, Dout, -1.0, Din)) / nelec < 1e-5) relmax = 0.5;
        double r = relax;
        if (fd > fd_old && std::fabs(fd - fd_old) > 1e-9) {
            r = relax / 4.0;
            relax *= 0.8;
        }
        if (fd < fd_old) relax *= 1.5;
        relax = std::min(relax, relmax);
        return axpby(r, Dout, 1.0 - r, Din);
    }
};

Mat density_from(const Ortho& X, const Mat& Fp, int nocc, std::vector<double>& eps, double& t_diag)
{
    const auto t0 = clk::now();
    Mat Cp;
    sym_eigen(Fp, eps, Cp);
    t_diag += secs(t0, clk::now());
    const int n = X.V.n, m = X.m;
    // C = V C' (n x m), D = 2 sum_{occ} C C^T
    Mat D(n);
    std::vector<double> col(n);
    for (int k = 0; k < nocc; ++k) {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += X.V(i, j) * Cp(j, k);
            col[i] = s;
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) D(i, j) += 2.0 * col[i] * col[j];
    }
    return D;
}

// D = V C' diag(occ) C'^T V^T over the orthonormal eigenvectors C' of F'.
Mat density_occ(const Ortho& X, const Mat& Cp, const std::vector<double>& occ)
{
    const int n = X.V.n, m = X.m;
    Mat D(n);
    std::vector<double> col(n);
    for (size_t k = 0; k < occ.size(); ++k) {
        if (occ[k] == 0.0) continue;
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += X.V(i, j) * Cp(j, (int)k);
            col[i] = s;
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) D(i, j) += occ[k] * col[i] * col[j];
    }
    return D;
}

// Aufbau with the open shell averaged: orbitals within 1e-4 Eh of each other share the electrons
// left for their group equally (spherical, fractional-occupation atom).
std::vector<double> averaged_occupations(const std::vector<double>& eps, double nelec)
{
    std::vector<double> occ(eps.size(), 0.0);
    double left = nelec;
    for (size_t i = 0; i < eps.size() && left > 1e-12;) {
        size_t j = i;
        while (j < eps.size() && std::fabs(eps[j] - eps[i]) < 1e-4) ++j;
        const double put = std::min(left, 2.0 * (double)(j - i));
        for (size_t k = i; k < j; ++k) occ[k] = put / (double)(j - i);
        left -= put;
        i = j;
    }
    return occ;
}

// Run fn(m) for m in [0, n) on a small pool of threads (the molecules or atoms are independent).
template <typename Fn>
void parallel_for(int n, int nthreads, Fn&& fn)
{
    nthreads = std::max(1, std::min(nthreads, n));
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errs(nthreads);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t] {
            try {
                for (int m = next++; m < n; m = next++) fn(m);
            } catch (...) {
                errs[t] = std::current_exception();
            }
        });
    for (auto& th : pool) th.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

int host_threads() { return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// Spherically averaged neutral-atom densities, each atom in its own shells: fractional-occupation
// RHF (averaged open shell), damped fixed point.  All atoms share one batch handle, so every
// iteration is one batched J/K build for all of them.
std::vector<Mat> atomic_densities(const std::vector<Basis>& ab, const std::vector<Atom>& at, const qcjk_opts& o)
{
    const int na = (int)ab.size();
    std::vector<Mat> D(na);
    if (na == 0) return D;
    std::vector<qcjk_basis> views(na);
    for (int m = 0; m < na; ++m) views[m] = ab[m].view();
    qcjk_handle* h = nullptr;
    int rc = qcjk_create_batch(views.data(), na, &o, &h);
    if (rc != QCJK_OK) throw EngineError("qcjk_create_batch", rc);
    struct Guard {
        qcjk_handle* h;
        ~Guard() { qcjk_destroy(h); }
    } guard{h};
    std::vector<int32_t> nbf(na);
    std::vector<int64_t> off(na);
    int64_t nelem = 0;
    rc = qcjk_layout(h, nullptr, nbf.data(), off.data(), &nelem);
    if (rc != QCJK_OK) throw EngineError("qcjk_layout", rc);
    std::vector<Mat> hcore(na);
    std::vector<Ortho> X(na);
    for (int m = 0; m < na; ++m) {
        const int n = nbf[m];
        Mat S(n), T(n), V(n);
        rc = qcjk_one_electron_mol(h, m, 1, &at[m].Z, at[m].R, S.a.data(), T.a.data(), V.a.data());
        if (rc != QCJK_OK) throw EngineError("qcjk_one_electron_mol", rc);
        hcore[m] = axpby(1.0, T, 1.0, V);
        X[m].set(S);
    }
    parallel_for(na, host_threads(), [&](int m) {
        std::vector<double> eps;
        Mat Cp;
        sym_eigen(X[m].to_ortho(hcore[m]), eps, Cp);
        D[m] = density_occ(X[m], Cp, averaged_occupations(eps, at[m].Z));
    });
    std::vector<char> done(na, 0);
    std::vector<double> Dp(nelem), Jp(nelem), Kp(nelem);
    for (int it = 0; it < 100; ++it) {
        for (int m = 0; m < na; ++m) std::memcpy(Dp.data() + off[m], D[m].a.data(), D[m].a.size() * sizeof(double));
        rc = qcjk_build(h, Dp.data(), Jp.data(), Kp.data(), QCJK_J | QCJK_K | QCJK_GRAPH, nullptr);
        if (rc != QCJK_OK) throw EngineError("qcjk_build", rc);
        parallel_for(na, host_threads(), [&](int m) {
            if (done[m]) return;
            const int n = nbf[m];
            Mat F = hcore[m];
            for (int k = 0; k < n * n; ++k) F.a[k] += Jp[off[m] + k] - 0.5 * Kp[off[m] + k];
            std::vector<double> eps;
            Mat Cp;
            sym_eigen(X[m].to_ortho(F), eps, Cp);
            const Mat Dn = density_occ(X[m], Cp, averaged_occupations(eps, at[m].Z));
            const double change = frob(axpby(1.0, Dn, -1.0, D[m]));
            D[m] = axpby(0.5, Dn, 0.5, D[m]);
            if (change < 1e-6) done[m] = 1;  // a seed: no need to go further
        });
        bool all = true;
        for (char c : done) all = all && c;
        if (all) break;
    }
    return D;
}

}  // namespace

std::vector<Mat> sad_densities(const std::vector<const Basis*>& bases, const std::vector<std::vector<Atom>>& atoms,
                               const qcjk_opts& o)
{
    // (atom signature, density), kept for the process: an element with the same shells always gets
    // the same atomic density, as from a tabulated SAD database
    static std::mutex cache_mu;
    static std::vector<std::pair<std::vector<double>, Mat>> cache;
    struct Place {
        int mol;
        std::vector<int> idx;      // molecular basis-function index of each atomic one
        std::vector<double> sig;
        int todo = -1;             // index into the batch of atoms still to compute
        Mat Da;
    };
    std::vector<Place> places;
    std::vector<Basis> todo_basis;
    std::vector<Atom> todo_atom;
    std::vector<std::vector<double>> todo_sig;
    for (size_t mi = 0; mi < bases.size(); ++mi) {
        const Basis& b = *bases[mi];
        std::vector<int> bf0(b.shells.size() + 1, 0);
        for (size_t s = 0; s < b.shells.size(); ++s) bf0[s + 1] = bf0[s] + (b.shells[s].L + 1) * (b.shells[s].L + 2) / 2;
        for (const Atom& a : atoms[mi]) {
            Place pl;
            pl.mol = (int)mi;
            Basis ab;
            for (size_t s = 0; s < b.shells.size(); ++s) {
                const Shell& sh = b.shells[s];
                const double dx = sh.R[0] - a.R[0], dy = sh.R[1] - a.R[1], dz = sh.R[2] - a.R[2];
                if (dx * dx + dy * dy + dz * dz > 1e-16) continue;
                ab.shells.push_back(sh);
                for (int k = bf0[s]; k < bf0[s + 1]; ++k) pl.idx.push_back(k);
            }
            if (ab.shells.empty()) continue;
            // same element with the same shells (translation-invariant): one atomic density
            pl.sig = {a.Z};
            for (const Shell& sh : ab.shells) {
                pl.sig.push_back(sh.L + 100.0 * sh.contracted);
                pl.sig.insert(pl.sig.end(), sh.exps.begin(), sh.exps.end());
                pl.sig.insert(pl.sig.end(), sh.coefs.begin(), sh.coefs.end());
            }
            {
                std::lock_guard<std::mutex> lk(cache_mu);
                for (const auto& c : cache)
                    if (c.first == pl.sig) pl.Da = c.second;
            }
            if (pl.Da.n == 0) {
                for (size_t t = 0; t < todo_sig.size() && pl.todo < 0; ++t)
                    if (todo_sig[t] == pl.sig) pl.todo = (int)t;
                if (pl.todo < 0) {
                    pl.todo = (int)todo_sig.size();
                    ab.finalize();
                    Atom centred = a;
                    todo_basis.push_back(std::move(ab));
                    todo_atom.push_back(centred);
                    todo_sig.push_back(pl.sig);
                }
            }
            places.push_back(std::move(pl));
        }
    }
    const std::vector<Mat> fresh = atomic_densities(todo_basis, todo_atom, o);
    {
        std::lock_guard<std::mutex> lk(cache_mu);
        for (size_t t = 0; t < fresh.size(); ++t) cache.emplace_back(todo_sig[t], fresh[t]);
    }
    std::vector<Mat> out;
    for (const Basis* b : bases) out.emplace_back(b->nbf);
    for (Place& pl : places) {
        const Mat& Da = pl.todo >= 0 ? fresh[pl.todo] : pl.Da;
        Mat& D = out[pl.mol];
        for (size_t i = 0; i < pl.idx.size(); ++i)
            for (size_t j = 0; j < pl.idx.size(); ++j) D(pl.idx[i], pl.idx[j]) = Da((int)i, (int)j);
    }
    return out;
}

Mat sad_density(const Basis& b, const std::vector<Atom>& atoms, const qcjk_opts& o)
{
    return sad_densities({&b}, {atoms}, o)[0];
}

double nuclear_repulsion(const std::vector<Atom>& atoms)
{
    double e = 0.0;
    for (size_t i = 0; i < atoms.size(); ++i)
        for (size_t j = 0; j < i; ++j) {
            const double dx = atoms[i].R[0] - atoms[j].R[0], dy = atoms[i].R[1] - atoms[j].R[1],
                         dz = atoms[i].R[2] - atoms[j].R[2];
            e += atoms[i].Z * atoms[j].Z / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

void one_electron(GpuHF_IBS& eng, const std::vector<Atom>& atoms, Mat& S, Mat& h)
{
    const int n = eng.nbf();
    std::vector<double> Z(atoms.size()), R(3 * atoms.size());
    for (size_t i = 0; i < atoms.size(); ++i) {
        Z[i] = atoms[i].Z;
        for (int x = 0; x < 3; ++x) R[3 * i + x] = atoms[i].R[x];
    }
    S = Mat(n);
    Mat T(n), V(n);
    const int rc = qcjk_one_electron(eng.handle(), (int)atoms.size(), Z.data(), R.data(), S.a.data(), T.a.data(), V.a.data());
    if (rc != QCJK_OK) throw EngineError("qcjk_one_electron", rc);
    h = axpby(1.0, T, 1.0, V);
}

// One closed-shell RHF run after its set-up: the part of the loop that follows each Fock build.
// Shared by rhf() (one molecule through the HF faces) and rhf_batch() (many molecules, one batched
// build per iteration), so both run the same recipe.
struct RHFRun {
    const SCFParams* par = nullptr;
    int n = 0, nocc = 0;
    Mat h;
    Ortho X;
    DIIS diis{4, 0.0, 0.0, {}, {}};
    Mat D;
    double E_old = 0.0, fd_old = 1e10;
    LinearMixer mixer;
    bool done = false;
    SCFResult res;

    void setup(const Mat& S, const Mat& h1, const std::vector<Atom>& atoms, int nelec, const SCFParams& p,
               const Mat* D0)
    {
        par = &p;
        if (nelec <= 0 || nelec % 2) throw std::runtime_error("rhf: closed-shell electron count required");
        nocc = nelec / 2;
        n = S.n;
        h = h1;
        X.set(S);
        if (nocc > X.m) throw std::runtime_error("rhf: more occupied orbitals than basis functions");
        res.e_nuc = nuclear_repulsion(atoms);
        double Ztot = 0.0;
        for (const Atom& a : atoms) Ztot += a.Z;
        diis = DIIS{p.nproj, p.fd_max > 0 ? p.fd_max : (double)((int)Ztot * (int)Ztot) * 0.1 / 32, p.sv_tol, {}, {}};
        D = D0 ? *D0 : density_from(X, X.to_ortho(h), nocc, res.eps, res.t_diag);
        if (D.n != n) throw std::runtime_error("rhf: seed density has the wrong dimension");
    }

    // F = h + J - 1/2 K was built from D.  Energy, convergence test, DIIS, new D.  Returns done.
    bool after_build(int it, const Mat& F, const Mat& J, const Mat& K, double t_elapsed)
    {
        res.e_one = dot(D, h);
        res.e_coul = 0.5 * dot(D, J);
        res.e_exch = -0.25 * dot(D, K);
        const double E = res.e_one + res.e_coul + res.e_exch + res.e_nuc;
        res.energy = E;
        res.energy_history.push_back(E);
        res.time_history.push_back(t_elapsed);
        res.iterations = it;

        const Mat Fp = X.to_ortho(F);
        const Mat Dp = X.density_to_ortho(D);
        const Mat Ep = axpby(1.0, Fp * Dp, -1.0, Dp * Fp);
        const double fd = frob(Ep);
        const double dE = it > 1 ? std::fabs(E - E_old) : 1e300;
        if (par->verbose)
            std::fprintf(stderr, "  it %3d  E = %.12f  dE = %.3e  [F,D] = %.3e\n", it, E, it > 1 ? E - E_old : 0.0, fd);
        if (dE < par->e_tol && fd < par->fd_tol) {
            res.converged = true;
            res.D = D;
            done = true;
            return true;
        }
        E_old = E;
        const Mat Fx = diis.next(Fp, Ep, fd);
        D = mixer.step(D, density_from(X, Fx, nocc, res.eps, res.t_diag), fd, fd_old, 2.0 * nocc);
        fd_old = fd;
        res.D = D;
        return false;
    }
};

SCFResult rhf(GpuHF_IBS& eng, const std::vector<Atom>& atoms, int nelec, const SCFParams& par, const Mat* D0,
              const Basis* basis_for_sad)
{
    const auto t_start = clk::now();
    const int n = eng.nbf();
    Mat S, h;
    one_electron(eng, atoms, S, h);
    Mat sad;
    if (!D0 && par.guess == 1) {
        if (!basis_for_sad) throw std::runtime_error("rhf: the SAD seed needs the basis");
        qcjk_opts o{};
        sad = sad_density(*basis_for_sad, atoms, o);
        D0 = &sad;
    }
    RHFRun run;
    run.setup(S, h, atoms, nelec, par, D0);
    run.res.t_setup = secs(t_start, clk::now());

    // The Fock operator terms as the reference's Hamiltonian holds them: the two HF terms reach the
    // basis through the WholeSystemFock face (J first, then K, same density -> one fused build).
    const WholeSystemFock_IBS<double>& whole = eng;
    Mat Dprev, Jprev, Kprev;  // incremental builds: J[D_k] = J[D_(k-1)] + J[D_k - D_(k-1)], same for K
    for (int it = 1; it <= par.max_iter; ++it) {
        const auto tj = clk::now();
        const bool delta = par.incremental > 0 && it > 1 && (it - 1) % par.incremental != 0;
        Mat Dao(n);
        whole.AddAODensity(Dao, delta ? axpby(1.0, run.D, -1.0, Dprev) : run.D);
        Mat J = whole.MakeAOFock(Dao, false);
        Mat K = whole.MakeAOFock(Dao, true);
        if (delta) {
            J += Jprev;
            K += Kprev;
        }
        if (par.incremental > 0) {
            Dprev = run.D;
            Jprev = J;
            Kprev = K;
        }
        run.res.t_jk += secs(tj, clk::now());
        Mat F = run.h;
        whole.SliceAOFock(F, J);
        Mat mhK = K;
        mhK *= -0.5;
        whole.SliceAOFock(F, mhK);
        if (run.after_build(it, F, J, K, secs(t_start, clk::now()))) break;
    }
    SCFResult res = std::move(run.res);
    res.t_total = secs(t_start, clk::now());
    res.quartets = eng.quartets_total();
    res.quartets_fp32 = eng.quartets_fp32_total();
    return res;
}

BatchSCFResult rhf_batch(const std::vector<const Basis*>& bases, const std::vector<std::vector<Atom>>& atoms,
                         const std::vector<int>& nelec, const SCFParams& par, const qcjk_opts& opts, unsigned jk_flags,
                         int nthreads, const std::vector<Mat>* seeds)
{
    const auto t_start = clk::now();
    const int nmol = (int)bases.size();
    if (nmol == 0 || (int)atoms.size() != nmol || (int)nelec.size() != nmol)
        throw std::runtime_error("rhf_batch: bases, atoms and electron counts differ in length");
    BatchSCFResult out;
    std::vector<qcjk_basis> views(nmol);
    for (int m = 0; m < nmol; ++m) views[m] = bases[m]->view();
    qcjk_handle* h = nullptr;
    int rc = qcjk_create_batch(views.data(), nmol, &opts, &h);
    if (rc != QCJK_OK) throw EngineError("qcjk_create_batch", rc);
    struct Guard {
        qcjk_handle* h;
        ~Guard() { qcjk_destroy(h); }
    } guard{h};
    std::vector<int32_t> nbf(nmol);
    std::vector<int64_t> off(nmol);
    int64_t nelem = 0;
    rc = qcjk_layout(h, nullptr, nbf.data(), off.data(), &nelem);
    if (rc != QCJK_OK) throw EngineError("qcjk_layout", rc);

    // per-molecule work on a small thread pool (molecules are independent)
    auto parallel = [&](auto&& fn) { parallel_for(nmol, nthreads, fn); };

    std::vector<RHFRun> runs(nmol);
    std::vector<Mat> S(nmol), h1(nmol);
    for (int m = 0; m < nmol; ++m) {  // one-electron integrals on the device, one molecule at a time
        const int n = nbf[m];
        std::vector<double> Z(atoms[m].size()), R(3 * atoms[m].size());
        for (size_t i = 0; i < atoms[m].size(); ++i) {
            Z[i] = atoms[m][i].Z;
            for (int x = 0; x < 3; ++x) R[3 * i + x] = atoms[m][i].R[x];
        }
        S[m] = Mat(n);
        Mat T(n), V(n);
        rc = qcjk_one_electron_mol(h, m, (int)Z.size(), Z.data(), R.data(), S[m].a.data(), T.a.data(), V.a.data());
        if (rc != QCJK_OK) throw EngineError("qcjk_one_electron_mol", rc);
        h1[m] = axpby(1.0, T, 1.0, V);
    }
    std::vector<Mat> sad(nmol);
    if (!seeds && par.guess == 1) {
        qcjk_opts ao{};
        ao.device = opts.device;
        sad = sad_densities(bases, atoms, ao);
        if (par.verbose) std::fprintf(stderr, "batch SAD seeds: %d molecules, %.2f s\n", nmol, secs(t_start, clk::now()));
    }
    SCFParams quiet = par;  // per-molecule iteration lines would be nmol per iteration
    quiet.verbose = false;
    parallel([&](int m) {
        const Mat* d0 = seeds ? &(*seeds)[m] : (par.guess == 1 ? &sad[m] : nullptr);
        runs[m].setup(S[m], h1[m], atoms[m], nelec[m], quiet, d0);
    });
    out.t_setup = secs(t_start, clk::now());

    std::vector<double> Dp(nelem), Jp(nelem), Kp(nelem);
    for (int it = 1; it <= par.max_iter; ++it) {
        const auto tj = clk::now();
        for (int m = 0; m < nmol; ++m) std::memcpy(Dp.data() + off[m], runs[m].D.a.data(), runs[m].D.a.size() * sizeof(double));
        qcjk_stats st{};
        rc = qcjk_build(h, Dp.data(), Jp.data(), Kp.data(), jk_flags | QCJK_GRAPH, &st);
        if (rc != QCJK_OK) throw EngineError("qcjk_build", rc);
        out.t_jk += secs(tj, clk::now());
        out.jk_device_ms += st.device_ms;
        out.quartets += st.quartets_computed;
        out.quartets_fp32 += st.quartets_fp32;
        out.iterations = it;
        const auto th = clk::now();
        const double t_now = secs(t_start, th);
        parallel([&](int m) {
            RHFRun& r = runs[m];
            if (r.done) return;
            const int n = nbf[m];
            Mat J(n), K(n);
            std::memcpy(J.a.data(), Jp.data() + off[m], J.a.size() * sizeof(double));
            std::memcpy(K.a.data(), Kp.data() + off[m], K.a.size() * sizeof(double));
            Mat F = r.h;
            F += J;
            Mat mhK = K;
            mhK *= -0.5;
            F += mhK;
            r.after_build(it, F, J, K, t_now);
        });
        out.t_host += secs(th, clk::now());
        int ndone = 0;
        for (const RHFRun& r : runs) ndone += r.done ? 1 : 0;
        if (par.verbose)  // one progress line per batched iteration
            std::fprintf(stderr, "batch iteration %d: %d/%d converged, %.2f s\n", it, ndone, nmol,
                         secs(t_start, clk::now()));
        if (ndone == nmol) break;
    }
    out.t_total = secs(t_start, clk::now());
    out.mols.reserve(nmol);
    for (RHFRun& r : runs) out.mols.push_back(std::move(r.res));
    return out;
}

}  // namespace qchem_amd
