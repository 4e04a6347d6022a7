// linalg.cpp -- see linalg.hpp.
#include "linalg.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace qchem_amd {

Mat& Mat::operator+=(const Mat& o)
{
    for (size_t i = 0; i < a.size(); ++i) a[i] += o.a[i];
    return *this;
}

Mat& Mat::operator*=(double s)
{
    for (double& v : a) v *= s;
    return *this;
}

Mat operator*(const Mat& x, const Mat& y)
{
    const int n = x.n;
    Mat r(n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double xik = x(i, k);
            if (xik == 0.0) continue;
            const double* yr = &y.a[(size_t)k * n];
            double* rr = &r.a[(size_t)i * n];
            for (int j = 0; j < n; ++j) rr[j] += xik * yr[j];
        }
    return r;
}

Mat transpose(const Mat& x)
{
    Mat r(x.n);
    for (int i = 0; i < x.n; ++i)
        for (int j = 0; j < x.n; ++j) r(j, i) = x(i, j);
    return r;
}

Mat axpby(double a, const Mat& x, double b, const Mat& y)
{
    Mat r(x.n);
    for (size_t i = 0; i < r.a.size(); ++i) r.a[i] = a * x.a[i] + b * y.a[i];
    return r;
}

double dot(const Mat& x, const Mat& y)
{
    double s = 0.0;
    for (size_t i = 0; i < x.a.size(); ++i) s += x.a[i] * y.a[i];
    return s;
}

double frob(const Mat& x) { return std::sqrt(dot(x, x)); }

// Householder reduction to tridiagonal form; z receives the accumulated transformation.
static void tridiagonalise(std::vector<double>& z, int n, std::vector<double>& d, std::vector<double>& e)
{
    auto Z = [&](int i, int j) -> double& { return z[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            double scale = 0.0;
            for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
            if (scale == 0.0) {
                e[i] = Z(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    Z(i, k) /= scale;
                    h += Z(i, k) * Z(i, k);
                }
                double f = Z(i, l);
                double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                Z(i, l) = f - g;
                // p = A u / h over the lower triangle, row by row (each row read once, contiguously)
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    Z(j, i) = Z(i, j) / h;
                    e[j] = 0.0;
                }
                const double* u = &Z(i, 0);
                for (int j = 0; j <= l; ++j) {
                    const double* row = &Z(j, 0);
                    const double uj = u[j];
                    double acc = row[j] * uj;
                    for (int k = 0; k < j; ++k) {
                        acc += row[k] * u[k];
                        e[k] += row[k] * uj;
                    }
                    e[j] += acc;
                }
                for (int j = 0; j <= l; ++j) {
                    e[j] /= h;
                    f += e[j] * Z(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = Z(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) Z(j, k) -= f * e[k] + g * Z(i, k);
                }
            }
        } else {
            e[i] = Z(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int l = i - 1;
        if (d[i] != 0.0 && l >= 0) {
            // g_j = sum_k Z(i,k) Z(k,j), then Z(k,j) -= g_j Z(k,i): columns j are independent, so
            // both passes run along rows
            std::vector<double> g(l + 1, 0.0);
            for (int k = 0; k <= l; ++k) {
                const double zik = Z(i, k);
                const double* row = &Z(k, 0);
                for (int j = 0; j <= l; ++j) g[j] += zik * row[j];
            }
            for (int k = 0; k <= l; ++k) {
                const double zki = Z(k, i);
                double* row = &Z(k, 0);
                for (int j = 0; j <= l; ++j) row[j] -= g[j] * zki;
            }
        }
        d[i] = Z(i, i);
        Z(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0.0;
    }
}

// Implicit QL iteration on the tridiagonal (d, e); rotations accumulated into the ROWS of zt (the
// transpose of the eigenvector matrix), so each rotation touches two contiguous rows.
static void ql_implicit(std::vector<double>& d, std::vector<double>& e, int n, std::vector<double>& zt)
{
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 + 2.2e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) throw std::runtime_error("sym_eigen: QL iteration did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    e[i + 1] = r = std::hypot(f, g);
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    double* zi = &zt[(size_t)i * n];
                    double* zj = &zt[(size_t)(i + 1) * n];
                    for (int k = 0; k < n; ++k) {
                        f = zj[k];
                        zj[k] = s * zi[k] + c * f;
                        zi[k] = c * zi[k] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

void sym_eigen(const Mat& A, std::vector<double>& w, Mat& V)
{
    const int n = A.n;
    std::vector<double> z(A.a), d(n), e(n);
    // symmetrise the input (use the mean of the two triangles)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double m = 0.5 * (z[(size_t)i * n + j] + z[(size_t)j * n + i]);
            z[(size_t)i * n + j] = z[(size_t)j * n + i] = m;
        }
    if (n > 0) {
        tridiagonalise(z, n, d, e);
        std::vector<double> zt((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) zt[(size_t)j * n + i] = z[(size_t)i * n + j];
        ql_implicit(d, e, n, zt);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) z[(size_t)i * n + j] = zt[(size_t)j * n + i];
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    V = Mat(n);
    for (int k = 0; k < n; ++k) {
        w[k] = d[idx[k]];
        for (int i = 0; i < n; ++i) V(i, k) = z[(size_t)i * n + idx[k]];
    }
}

bool solve(std::vector<double> A, std::vector<double> b, int n, std::vector<double>& x)
{
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
            b[i] -= f * b[k];
        }
    }
    x.assign(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * x[j];
        x[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

}  // namespace qchem_amd
