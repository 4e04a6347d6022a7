// linalg.hpp -- the small dense linear algebra the RHF driver needs (no LAPACK in this image).
//
// Stands in for the reference's LAPACK-through-Blaze calls (LASolver/Internal/LASolverLapack.C:10-128,
// Math/DIIS.C:37-51): symmetric eigendecomposition (Householder tridiagonalisation followed by the
// implicit QL iteration with Wilkinson-style shifts -- Wilkinson & Reinsch, Handbook for Automatic
// Computation II (1971), procedures tred2/tql2), and a pivoted Gaussian solve for the DIIS system.
#pragma once

#include <cstddef>
#include <vector>

namespace qchem_amd {

// Dense row-major square matrix.
struct Mat {
    int n = 0;
    std::vector<double> a;
    Mat() = default;
    explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
    double& operator()(int i, int j) { return a[(size_t)i * n + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * n + j]; }
    Mat& operator+=(const Mat& o);
    Mat& operator*=(double s);
};

Mat operator*(const Mat& x, const Mat& y);          // product
Mat transpose(const Mat& x);
Mat axpby(double a, const Mat& x, double b, const Mat& y);  // a x + b y
double dot(const Mat& x, const Mat& y);              // sum_ij x_ij y_ij
double frob(const Mat& x);

// Eigen-decomposition of a symmetric matrix: w ascending, V columns are eigenvectors.
void sym_eigen(const Mat& A, std::vector<double>& w, Mat& V);

// Solve A x = b (A general, small); returns false if singular.
bool solve(std::vector<double> A, std::vector<double> b, int n, std::vector<double>& x);

}  // namespace qchem_amd
