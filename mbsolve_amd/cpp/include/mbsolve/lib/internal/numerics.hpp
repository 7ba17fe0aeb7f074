/*
 * Host-side FP64 linear algebra for the solver set-up: small dense real and
 * complex matrices, the matrix exponential, a full-pivot LU solve and the
 * real block decomposition of a skew-symmetric matrix.  These replace the
 * Eigen calls of the reference set-up code (MatrixBase::exp,
 * FullPivLU::solve, RealSchur in algo_lindblad_cvr_rodr.hpp:104,185,200 and
 * algo_lindblad_reg_cayley.hpp:169,188).
 */
#ifndef MBSOLVE_LIB_INTERNAL_NUMERICS_H
#define MBSOLVE_LIB_INTERNAL_NUMERICS_H

#include <complex>
#include <vector>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {
namespace num {

/** Dense row-major matrix. */
template<typename T>
struct matrix {
    size_t rows = 0, cols = 0;
    std::vector<T> a;

    matrix() {}
    matrix(size_t r, size_t c) : rows(r), cols(c), a(r * c, T(0)) {}

    T& operator()(size_t i, size_t j) { return a[i * cols + j]; }
    const T& operator()(size_t i, size_t j) const { return a[i * cols + j]; }

    static matrix identity(size_t n)
    {
        matrix m(n, n);
        for (size_t i = 0; i < n; i++) {
            m(i, i) = T(1);
        }
        return m;
    }
};

typedef matrix<real> rmat;
typedef matrix<std::complex<real> > cmat;

template<typename T>
matrix<T> operator*(const matrix<T>& A, const matrix<T>& B)
{
    matrix<T> C(A.rows, B.cols);
    for (size_t i = 0; i < A.rows; i++) {
        for (size_t j = 0; j < B.cols; j++) {
            T acc = A(i, 0) * B(0, j);
            for (size_t k = 1; k < A.cols; k++) {
                acc += A(i, k) * B(k, j);
            }
            C(i, j) = acc;
        }
    }
    return C;
}

template<typename T>
matrix<T> operator+(const matrix<T>& A, const matrix<T>& B)
{
    matrix<T> C(A.rows, A.cols);
    for (size_t i = 0; i < A.a.size(); i++) {
        C.a[i] = A.a[i] + B.a[i];
    }
    return C;
}

template<typename T>
matrix<T> operator-(const matrix<T>& A, const matrix<T>& B)
{
    matrix<T> C(A.rows, A.cols);
    for (size_t i = 0; i < A.a.size(); i++) {
        C.a[i] = A.a[i] - B.a[i];
    }
    return C;
}

template<typename T>
matrix<T> scaled(const matrix<T>& A, T s)
{
    matrix<T> C(A.rows, A.cols);
    for (size_t i = 0; i < A.a.size(); i++) {
        C.a[i] = A.a[i] * s;
    }
    return C;
}

template<typename T>
matrix<T> transpose(const matrix<T>& A)
{
    matrix<T> C(A.cols, A.rows);
    for (size_t i = 0; i < A.rows; i++) {
        for (size_t j = 0; j < A.cols; j++) {
            C(j, i) = A(i, j);
        }
    }
    return C;
}

/** y = A x */
std::vector<real> matvec(const rmat& A, const std::vector<real>& x);

/** Maximum absolute column sum. */
real norm1(const rmat& A);

/** Solves A X = B (A square) by Gaussian elimination with partial
 * pivoting. */
rmat solve_partial_pivot(const rmat& A, const rmat& B);

/** Solves A x = b by LU decomposition with complete (full) pivoting;
 * singular directions get zero components. */
std::vector<real> solve_full_pivot(const rmat& A, const std::vector<real>& b);

/**
 * Matrix exponential by scaling and squaring with diagonal Pade
 * approximants of degree 3, 5, 7, 9 or 13 (N. J. Higham, "The scaling and
 * squaring method for the matrix exponential revisited", SIAM J. Matrix
 * Anal. Appl. 26 (2005) 1179-1193, Algorithm 2.3).
 */
rmat expm(const rmat& A);

/**
 * Real block decomposition of a skew-symmetric K x K matrix U:
 * U = Q T Q^T with Q orthogonal and T block diagonal; the first
 * 2 * theta.size() columns of Q form pairs (q1_j, q2_j) = (col 2j,
 * col 2j+1) with U q1 = theta_j q2 and U q2 = -theta_j q1, theta_j > 0
 * sorted in descending order; the remaining columns span the null space.
 *
 * Computed from the eigen-decomposition of the real symmetric 2K x 2K
 * matrix [[0, -U], [U, 0]] (the real form of the Hermitian matrix i U)
 * by the cyclic Jacobi method.
 */
struct skew_blocks {
    rmat Q;
    std::vector<real> theta;
};
skew_blocks skew_decompose(const rmat& U, real rel_tol = 1e-13);

/** Eigen-decomposition of a real symmetric matrix by cyclic Jacobi
 * rotations: A = V diag(w) V^T (w unsorted). */
void jacobi_eigen(const rmat& A, std::vector<real>& w, rmat& V);

} /* namespace num */
}

#endif
