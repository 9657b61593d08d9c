// Small dense complex factorisations done once per state on the host (D x D, D <= a few
// hundred): Schur form, eigenvectors of the no-jump generator Q, inverses.  They prepare the
// spectral data that the GPU preconditioners apply.
#pragma once
#include <complex>
#include <vector>

namespace rc {

using cd = std::complex<double>;
using Mat = std::vector<cd>;   // row-major n x n

// A = Z T Z^H with T upper triangular.  Returns false if QR iteration did not converge.
bool schur(int n, Mat& A_to_T, Mat& Z);
// Eigenvectors of upper-triangular T (columns of W, unit diagonal, W upper triangular).
void triangular_eigenvectors(int n, const Mat& T, Mat& W);
// Inverse of an upper-triangular matrix.
bool upper_inverse(int n, const Mat& U, Mat& Uinv);
// General inverse by LU with partial pivoting.
bool inverse(int n, const Mat& A, Mat& Ainv);
// Spectral norm (power iteration on A^H A).
double opnorm2(int n, const Mat& A);
// C = A * B
void matmul(int n, const Mat& A, const Mat& B, Mat& C);
Mat adjoint(int n, const Mat& A);

}  // namespace rc
