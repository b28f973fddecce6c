// Host eigensolvers of the small matrices the Krylov drivers produce: the symmetric tridiagonal of
// a Lanczos run and Krylov-Schur's projected matrix. Implicit QL with Wilkinson shifts (after
// Householder reduction for a dense matrix), the textbook algorithms; they replace the LAPACK
// dstevd / dstevx / dsyevd calls of the reference's include/ed/krylov/tridiag.h.
#pragma once

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

/// Eigenpairs of a dense symmetric m x m matrix (column-major).
TridiagEig symmetric_eig(std::vector<double> a, std::size_t m);

/// Eigenvalues of a dense Hermitian n x n matrix (column-major, re/im interleaved), ascending.
std::vector<double> hermitian_eigvals(const std::vector<double>& a, std::size_t n);

}  // namespace qsa
