/*
 * Host FP64 linear algebra used by the solver set-up (see numerics.hpp).
 */
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <mbsolve/lib/internal/numerics.hpp>

namespace mbsolve {
namespace num {

std::vector<real>
matvec(const rmat& A, const std::vector<real>& x)
{
    std::vector<real> y(A.rows, 0.0);
    for (size_t i = 0; i < A.rows; i++) {
        real acc = A(i, 0) * x[0];
        for (size_t k = 1; k < A.cols; k++) {
            acc += A(i, k) * x[k];
        }
        y[i] = acc;
    }
    return y;
}

real
norm1(const rmat& A)
{
    real best = 0.0;
    for (size_t j = 0; j < A.cols; j++) {
        real s = 0.0;
        for (size_t i = 0; i < A.rows; i++) {
            s += std::abs(A(i, j));
        }
        best = std::max(best, s);
    }
    return best;
}

rmat
solve_partial_pivot(const rmat& A_in, const rmat& B_in)
{
    const size_t n = A_in.rows;
    rmat A = A_in, B = B_in;
    for (size_t col = 0; col < n; col++) {
        size_t piv = col;
        for (size_t r = col + 1; r < n; r++) {
            if (std::abs(A(r, col)) > std::abs(A(piv, col))) {
                piv = r;
            }
        }
        if (A(piv, col) == 0.0) {
            throw std::runtime_error("singular matrix in linear solve");
        }
        if (piv != col) {
            for (size_t k = 0; k < n; k++) {
                std::swap(A(col, k), A(piv, k));
            }
            for (size_t k = 0; k < B.cols; k++) {
                std::swap(B(col, k), B(piv, k));
            }
        }
        for (size_t r = col + 1; r < n; r++) {
            const real f = A(r, col) / A(col, col);
            if (f == 0.0) {
                continue;
            }
            for (size_t k = col; k < n; k++) {
                A(r, k) -= f * A(col, k);
            }
            for (size_t k = 0; k < B.cols; k++) {
                B(r, k) -= f * B(col, k);
            }
        }
    }
    rmat X(n, B.cols);
    for (size_t k = 0; k < B.cols; k++) {
        for (size_t ii = n; ii-- > 0;) {
            real acc = B(ii, k);
            for (size_t j = ii + 1; j < n; j++) {
                acc -= A(ii, j) * X(j, k);
            }
            X(ii, k) = acc / A(ii, ii);
        }
    }
    return X;
}

std::vector<real>
solve_full_pivot(const rmat& A_in, const std::vector<real>& b_in)
{
    const size_t n = A_in.rows;
    rmat A = A_in;
    std::vector<real> b = b_in;
    std::vector<size_t> colperm(n);
    std::iota(colperm.begin(), colperm.end(), 0);
    real maxpiv = 0.0;
    size_t rank = n;
    for (size_t k = 0; k < n; k++) {
        size_t pr = k, pc = k;
        real best = -1.0;
        for (size_t i = k; i < n; i++) {
            for (size_t j = k; j < n; j++) {
                if (std::abs(A(i, j)) > best) {
                    best = std::abs(A(i, j));
                    pr = i;
                    pc = j;
                }
            }
        }
        maxpiv = std::max(maxpiv, best);
        if (best <= maxpiv * n * 2.220446049250313e-16 || best == 0.0) {
            rank = k;
            break;
        }
        for (size_t j = 0; j < n; j++) {
            std::swap(A(k, j), A(pr, j));
        }
        std::swap(b[k], b[pr]);
        for (size_t i = 0; i < n; i++) {
            std::swap(A(i, k), A(i, pc));
        }
        std::swap(colperm[k], colperm[pc]);
        for (size_t i = k + 1; i < n; i++) {
            const real f = A(i, k) / A(k, k);
            for (size_t j = k; j < n; j++) {
                A(i, j) -= f * A(k, j);
            }
            b[i] -= f * b[k];
        }
    }
    std::vector<real> y(n, 0.0);
    for (size_t ii = rank; ii-- > 0;) {
        real acc = b[ii];
        for (size_t j = ii + 1; j < rank; j++) {
            acc -= A(ii, j) * y[j];
        }
        y[ii] = acc / A(ii, ii);
    }
    std::vector<real> x(n, 0.0);
    for (size_t k = 0; k < n; k++) {
        x[colperm[k]] = y[k];
    }
    return x;
}

namespace {

/* Pade coefficients b_0..b_m of the [m/m] approximant of exp. */
const double pade3[] = { 120.0, 60.0, 12.0, 1.0 };
const double pade5[] = { 30240.0, 15120.0, 3360.0, 420.0, 30.0, 1.0 };
const double pade7[] = { 17297280.0, 8648640.0, 1995840.0, 277200.0,
                         25200.0,    1512.0,    56.0,      1.0 };
const double pade9[] = { 17643225600.0, 8821612800.0, 2075673600.0,
                         302702400.0,   30270240.0,   2162160.0,
                         110880.0,      3960.0,       90.0,
                         1.0 };
const double pade13[] = { 64764752532480000.0,
                          32382376266240000.0,
                          7771770303897600.0,
                          1187353796428800.0,
                          129060195264000.0,
                          10559470521600.0,
                          670442572800.0,
                          33522128640.0,
                          1323241920.0,
                          40840800.0,
                          960960.0,
                          16380.0,
                          182.0,
                          1.0 };

/* Returns exp(A) from U (odd part) and V (even part): (V - U)^-1 (V + U). */
rmat pade_finish(const rmat& U, const rmat& V)
{
    return solve_partial_pivot(V - U, V + U);
}

rmat pade_low(const rmat& A, const double* b, int m)
{
    const size_t n = A.rows;
    const rmat I = rmat::identity(n);
    const rmat A2 = A * A;
    /* powers of A^2 */
    std::vector<rmat> P;
    P.push_back(I);
    for (int k = 1; 2 * k <= m; k++) {
        P.push_back(P.back() * A2);
    }
    rmat Uo(n, n), V(n, n);
    for (int k = 0; 2 * k + 1 <= m; k++) {
        Uo = Uo + scaled(P[k], b[2 * k + 1]);
    }
    for (int k = 0; 2 * k <= m; k++) {
        V = V + scaled(P[k], b[2 * k]);
    }
    return pade_finish(A * Uo, V);
}

rmat pade13_eval(const rmat& A)
{
    const size_t n = A.rows;
    const double* b = pade13;
    const rmat I = rmat::identity(n);
    const rmat A2 = A * A, A4 = A2 * A2, A6 = A4 * A2;
    rmat t1 = scaled(A6, b[13]) + scaled(A4, b[11]) + scaled(A2, b[9]);
    rmat t2 = scaled(A6, b[7]) + scaled(A4, b[5]) + scaled(A2, b[3]) +
        scaled(I, b[1]);
    const rmat U = A * (A6 * t1 + t2);
    rmat t3 = scaled(A6, b[12]) + scaled(A4, b[10]) + scaled(A2, b[8]);
    rmat t4 = scaled(A6, b[6]) + scaled(A4, b[4]) + scaled(A2, b[2]) +
        scaled(I, b[0]);
    const rmat V = A6 * t3 + t4;
    return pade_finish(U, V);
}

} /* namespace */

rmat
expm(const rmat& A)
{
    if (A.rows != A.cols) {
        throw std::invalid_argument("expm: matrix must be square");
    }
    const real nrm = norm1(A);
    /* theta_m of Higham (2005), Table 2.3 */
    if (nrm <= 1.495585217958292e-2) {
        return pade_low(A, pade3, 3);
    }
    if (nrm <= 2.539398330063230e-1) {
        return pade_low(A, pade5, 5);
    }
    if (nrm <= 9.504178996162932e-1) {
        return pade_low(A, pade7, 7);
    }
    if (nrm <= 2.097847961257068e0) {
        return pade_low(A, pade9, 9);
    }
    const real theta13 = 5.371920351148152;
    int s = 0;
    if (nrm > theta13) {
        s = (int)std::ceil(std::log2(nrm / theta13));
    }
    rmat X = pade13_eval(scaled(A, std::ldexp(1.0, -s)));
    for (int i = 0; i < s; i++) {
        X = X * X;
    }
    return X;
}

void
jacobi_eigen(const rmat& A_in, std::vector<real>& w, rmat& V)
{
    const size_t n = A_in.rows;
    rmat A = A_in;
    V = rmat::identity(n);
    for (int sweep = 0; sweep < 100; sweep++) {
        real off = 0.0, diag = 0.0;
        for (size_t p = 0; p < n; p++) {
            diag += A(p, p) * A(p, p);
            for (size_t q = p + 1; q < n; q++) {
                off += A(p, q) * A(p, q);
            }
        }
        if (off == 0.0 || off <= 1e-36 * (diag + off)) {
            break;
        }
        for (size_t p = 0; p < n; p++) {
            for (size_t q = p + 1; q < n; q++) {
                const real apq = A(p, q);
                if (apq == 0.0) {
                    continue;
                }
                const real tau = (A(q, q) - A(p, p)) / (2.0 * apq);
                const real t = (tau >= 0 ? 1.0 : -1.0) /
                    (std::abs(tau) + std::sqrt(1.0 + tau * tau));
                const real c = 1.0 / std::sqrt(1.0 + t * t);
                const real s = t * c;
                for (size_t k = 0; k < n; k++) {
                    const real akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (size_t k = 0; k < n; k++) {
                    const real apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                for (size_t k = 0; k < n; k++) {
                    const real vkp = V(k, p), vkq = V(k, q);
                    V(k, p) = c * vkp - s * vkq;
                    V(k, q) = s * vkp + c * vkq;
                }
            }
        }
    }
    w.resize(n);
    for (size_t i = 0; i < n; i++) {
        w[i] = A(i, i);
    }
}

namespace {

real dot(const std::vector<real>& a, const std::vector<real>& b)
{
    real s = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        s += a[i] * b[i];
    }
    return s;
}

/* Orthogonalises v against the columns in `basis` (twice, modified
 * Gram-Schmidt) and returns the remaining norm. */
real orthogonalize(std::vector<real>& v,
                   const std::vector<std::vector<real> >& basis)
{
    for (int pass = 0; pass < 2; pass++) {
        for (const auto& q : basis) {
            const real c = dot(v, q);
            for (size_t i = 0; i < v.size(); i++) {
                v[i] -= c * q[i];
            }
        }
    }
    return std::sqrt(dot(v, v));
}

} /* namespace */

skew_blocks
skew_decompose(const rmat& U, real rel_tol)
{
    const size_t K = U.rows;
    real unorm = 0.0;
    for (real x : U.a) {
        unorm = std::max(unorm, std::abs(x));
    }
    skew_blocks out;
    out.Q = rmat(K, K);
    if (unorm == 0.0) {
        out.Q = rmat::identity(K);
        return out;
    }
    /* real symmetric form of the Hermitian matrix iU */
    rmat S(2 * K, 2 * K);
    for (size_t i = 0; i < K; i++) {
        for (size_t j = 0; j < K; j++) {
            S(i, K + j) = -U(i, j);
            S(K + i, j) = U(i, j);
        }
    }
    std::vector<real> w;
    rmat V;
    jacobi_eigen(S, w, V);
    std::vector<size_t> order(2 * K);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&w](size_t a, size_t b) { return w[a] > w[b]; });

    const real tol = rel_tol * unorm * K;
    std::vector<std::vector<real> > basis;
    std::vector<real> thetas;
    for (size_t oi = 0; oi < order.size(); oi++) {
        const size_t e = order[oi];
        if (w[e] <= tol) {
            break;
        }
        /* eigenvector (x; y) of [[0,-U],[U,0]]:  U x = w y, U y = -w x */
        std::vector<real> x(K), y(K);
        for (size_t i = 0; i < K; i++) {
            x[i] = V(i, e);
            y[i] = V(K + i, e);
        }
        const real xn = std::sqrt(dot(x, x));
        std::vector<real> q1 = x;
        real r = orthogonalize(q1, basis);
        if (r < 0.5 * xn || xn < 1e-3) {
            continue; /* plane already taken (doubled eigenvalue) */
        }
        for (auto& v : q1) {
            v /= r;
        }
        /* q2 = U q1 / |U q1| */
        std::vector<real> q2(K, 0.0);
        for (size_t i = 0; i < K; i++) {
            real acc = 0.0;
            for (size_t j = 0; j < K; j++) {
                acc += U(i, j) * q1[j];
            }
            q2[i] = acc;
        }
        std::vector<std::vector<real> > b2 = basis;
        b2.push_back(q1);
        const real r2 = orthogonalize(q2, b2);
        if (r2 <= tol) {
            continue;
        }
        for (auto& v : q2) {
            v /= r2;
        }
        basis.push_back(q1);
        basis.push_back(q2);
        /* theta = q2^T U q1 */
        real th = 0.0;
        for (size_t i = 0; i < K; i++) {
            real acc = 0.0;
            for (size_t j = 0; j < K; j++) {
                acc += U(i, j) * q1[j];
            }
            th += q2[i] * acc;
        }
        thetas.push_back(th);
        if (2 * thetas.size() + 1 > K) {
            break;
        }
    }
    /* null space: complete the basis with the near-zero eigenvectors, then
     * with unit vectors if rounding left gaps */
    for (size_t oi = 0; oi < order.size() && basis.size() < K; oi++) {
        const size_t e = order[oi];
        if (std::abs(w[e]) > tol) {
            continue;
        }
        for (int part = 0; part < 2 && basis.size() < K; part++) {
            std::vector<real> v(K);
            for (size_t i = 0; i < K; i++) {
                v[i] = V(part * K + i, e);
            }
            const real vn = std::sqrt(dot(v, v));
            if (vn < 1e-3) {
                continue;
            }
            real r = orthogonalize(v, basis);
            if (r < 0.5 * vn) {
                continue;
            }
            for (auto& x : v) {
                x /= r;
            }
            basis.push_back(v);
        }
    }
    for (size_t u = 0; u < K && basis.size() < K; u++) {
        std::vector<real> v(K, 0.0);
        v[u] = 1.0;
        real r = orthogonalize(v, basis);
        if (r > 0.5) {
            for (auto& x : v) {
                x /= r;
            }
            basis.push_back(v);
        }
    }
    if (basis.size() != K) {
        throw std::runtime_error("skew_decompose: basis incomplete");
    }
    /* sort planes by descending theta (basis pairs move together) */
    std::vector<size_t> porder(thetas.size());
    std::iota(porder.begin(), porder.end(), 0);
    std::stable_sort(porder.begin(), porder.end(), [&](size_t a, size_t b) {
        return thetas[a] > thetas[b];
    });
    out.theta.resize(thetas.size());
    for (size_t j = 0; j < porder.size(); j++) {
        out.theta[j] = thetas[porder[j]];
        for (size_t i = 0; i < K; i++) {
            out.Q(i, 2 * j) = basis[2 * porder[j]][i];
            out.Q(i, 2 * j + 1) = basis[2 * porder[j] + 1][i];
        }
    }
    for (size_t c = 2 * thetas.size(); c < K; c++) {
        for (size_t i = 0; i < K; i++) {
            out.Q(i, c) = basis[c][i];
        }
    }
    return out;
}

} /* namespace num */
}
