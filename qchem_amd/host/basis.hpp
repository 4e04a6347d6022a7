// basis.hpp -- molecular Cartesian Gaussian basis in the reference conventions, flattened for the
// engine's C ABI.
//
//   folded coefficient  contracted radial: c_i * N(alpha_i; L,0,0)  (GaussianRF.C:566-577, :512-521)
//                       single-exponent radial: 1                   (GaussianRF.C:557-564)
//   ns (per component)  1/sqrt(S_ii)                                (PGData.C:27-31)
//   component order     m outer, l inner, n = L-m-l                 (PG_Cart/Imp/IrrepBasisSet.C:46-54)
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

#include "qcjk.h"

namespace qchem_amd {

struct Shell {
    int L = 0;
    std::vector<double> exps, coefs;  // raw contraction coefficients
    double R[3] = {0, 0, 0};
    bool contracted = false;          // true: the contracted-radial constructor (norm folded)
};

inline double dfact_odd(int k)  // (2k-1)!!
{
    double r = 1.0;
    for (int i = 2 * k - 1; i > 1; i -= 2) r *= i;
    return r;
}

struct Basis {
    std::vector<Shell> shells;
    int nbf = 0;
    // flattened
    std::vector<int> l, np;
    std::vector<double> cen, exps, folded, ns;

    void finalize()
    {
        const double pi = 3.14159265358979323846;
        l.clear(); np.clear(); cen.clear(); exps.clear(); folded.clear(); ns.clear();
        nbf = 0;
        for (const Shell& s : shells) {
            if (s.exps.empty() || s.exps.size() != s.coefs.size()) throw std::runtime_error("Basis: bad shell");
            l.push_back(s.L);
            np.push_back((int)s.exps.size());
            for (int x = 0; x < 3; ++x) cen.push_back(s.R[x]);
            std::vector<double> c(s.exps.size());
            for (size_t i = 0; i < s.exps.size(); ++i) {
                const double a = s.exps[i];
                if (s.contracted) {
                    const double norm = std::sqrt(std::pow(4 * a, s.L) / (std::pow(pi / (2 * a), 1.5) * dfact_odd(s.L)));
                    c[i] = s.coefs[i] * norm;
                } else {
                    c[i] = 1.0;
                }
                exps.push_back(a);
                folded.push_back(c[i]);
            }
            for (int m = 0; m <= s.L; ++m)
                for (int ll = 0; ll <= s.L - m; ++ll) {
                    const int pw[3] = {s.L - m - ll, ll, m};
                    double S = 0.0;
                    for (size_t i = 0; i < s.exps.size(); ++i)
                        for (size_t j = 0; j < s.exps.size(); ++j) {
                            const double p = s.exps[i] + s.exps[j];
                            double f = std::pow(pi / p, 1.5);
                            for (int x = 0; x < 3; ++x) f *= dfact_odd(pw[x]) / std::pow(2 * p, pw[x]);
                            S += c[i] * c[j] * f;
                        }
                    ns.push_back(1.0 / std::sqrt(S));
                    ++nbf;
                }
        }
    }

    qcjk_basis view() const
    {
        qcjk_basis b{};
        b.nshell = (int)l.size();
        b.shell_l = l.data();
        b.shell_nprim = np.data();
        b.shell_center = cen.data();
        b.prim_exp = exps.data();
        b.prim_coef = folded.data();
        b.bf_norm = ns.data();
        b.order = 0;
        return b;
    }
};

}  // namespace qchem_amd
