// Host eigensolvers of the small matrices the Krylov drivers produce: the symmetric tridiagonal of
// a Lanczos run and Krylov-Schur's projected matrix. Implicit QL with Wilkinson shifts (after
// Householder reduction for a dense matrix), the textbook algorithms; they replace the LAPACK
// dstevd / dstevx / dsyevd calls of the reference's include/ed/krylov/tridiag.h.
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

namespace qsa {

/// Eigenvalues ascending; vectors column-major (column j belongs to values[j]).
struct TridiagEig {
    std::size_t m = 0;
    std::vector<double> values;
    std::vector<double> vectors;
    double z(std::size_t row, std::size_t col) const { return vectors[col * m + row]; }
    std::vector<double> weights() const;   ///< squared first components
};

/// The leading m x m block of the tridiagonal with diagonal alpha[0..m) and off-diagonal
/// beta[1..m) (beta[i] couples steps i-1 and i). Throws qsa::Error(kConvergence) on a
/// non-finite entry or when QL does not converge.
TridiagEig tridiag_eig(const std::vector<double>& alpha, const std::vector<double>& beta, std::size_t m,
                       bool vectors);

/// Ritz values (ascending) and squared first components of the m x m tridiagonal with diagonal
/// alpha[0..m) and off-diagonal beta_next[0..m-1) (beta_next[i] couples i and i+1): the same values
/// and weights as tridiag_eig(..., true), bit for bit, in O(m^2).
void tridiag_weights(const double* alpha, const double* beta_next, std::size_t m, std::vector<double>& values,
                     std::vector<double>& weights);

/// Eigenpairs of a dense symmetric m x m matrix (column-major).
TridiagEig symmetric_eig(std::vector<double> a, std::size_t m);

/// Eigenvalues of a dense Hermitian n x n matrix (column-major), ascending: Householder reduction
/// to a tridiagonal with complex reflectors, then implicit QL (the O(n^3) dense lane).
std::vector<double> hermitian_eigvals_householder(std::vector<std::complex<double>> a, std::size_t n);

/// Eigenvalues of a dense Hermitian n x n matrix (column-major, re/im interleaved), ascending.
std::vector<double> hermitian_eigvals(const std::vector<double>& a, std::size_t n);

}  // namespace qsa
