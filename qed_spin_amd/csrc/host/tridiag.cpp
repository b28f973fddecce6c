#include "tridiag.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <string>

#include "hip_util.h"

namespace qsa {

std::vector<double> TridiagEig::weights() const {
    std::vector<double> w(m);
    for (std::size_t j = 0; j < m; ++j) w[j] = vectors[j * m] * vectors[j * m];
    return w;
}

namespace {

// Implicit QL with Wilkinson shifts on the tridiagonal (d, e), e[i] coupling i and i+1, rotations
// accumulated into the column-major n x n matrix z (when non-null).
void ql_implicit(std::vector<double>& d, std::vector<double>& e, std::size_t n, double* z) {
    if (n == 0) return;
    e.resize(n, 0.0);
    e[n - 1] = 0.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (std::size_t l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            std::size_t m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                if (std::abs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 90) throw Error(kConvergence, "tridiagonal eigensolve: QL did not converge");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool underflow = false;
            for (std::size_t ii = m; ii-- > l;) {
                const double f = s * e[ii];
                const double b = c * e[ii];
                r = std::hypot(f, g);
                e[ii + 1] = r;
                if (r == 0.0) {
                    d[ii + 1] -= p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[ii + 1] - p;
                r = (d[ii] - g) * s + 2.0 * c * b;
                p = s * r;
                d[ii + 1] = g + p;
                g = c * r - b;
                if (z) {
                    double* zi = z + ii * n;
                    double* zj = z + (ii + 1) * n;
                    for (std::size_t k = 0; k < n; ++k) {
                        const double fk = zj[k];
                        zj[k] = s * zi[k] + c * fk;
                        zi[k] = c * zi[k] - s * fk;
                    }
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

TridiagEig sorted(std::vector<double>& d, std::vector<double>* z, std::size_t n) {
    TridiagEig t;
    t.m = n;
    std::vector<std::size_t> order(n);
    std::iota(order.begin(), order.end(), std::size_t{0});
    std::stable_sort(order.begin(), order.end(), [&d](std::size_t a, std::size_t b) { return d[a] < d[b]; });
    t.values.resize(n);
    for (std::size_t j = 0; j < n; ++j) t.values[j] = d[order[j]];
    if (z) {
        t.vectors.resize(n * n);
        for (std::size_t j = 0; j < n; ++j)
            std::copy(z->begin() + static_cast<std::ptrdiff_t>(order[j] * n),
                      z->begin() + static_cast<std::ptrdiff_t>((order[j] + 1) * n), t.vectors.begin() + j * n);
    }
    return t;
}

}  // namespace

TridiagEig tridiag_eig(const std::vector<double>& alpha, const std::vector<double>& beta, std::size_t m,
                       bool vectors) {
    if (m == 0) return TridiagEig{};
    std::vector<double> d(alpha.begin(), alpha.begin() + static_cast<std::ptrdiff_t>(m));
    std::vector<double> e(m, 0.0);
    for (std::size_t i = 0; i + 1 < m; ++i) e[i] = beta[i + 1];
    for (std::size_t i = 0; i < m; ++i)
        if (!std::isfinite(d[i]) || !std::isfinite(e[i]))
            throw Error(kConvergence, "tridiagonal eigensolve: non-finite entry at step " + std::to_string(i));
    std::vector<double> z;
    if (vectors) {
        z.assign(m * m, 0.0);
        for (std::size_t i = 0; i < m; ++i) z[i * m + i] = 1.0;
    }
    ql_implicit(d, e, m, vectors ? z.data() : nullptr);
    return sorted(d, vectors ? &z : nullptr, m);
}

TridiagEig symmetric_eig(std::vector<double> a, std::size_t n) {
    if (n == 0) return TridiagEig{};
    for (double x : a)
        if (!std::isfinite(x)) throw Error(kConvergence, "projected eigensolve: non-finite entry");
    // Householder reduction A = Q T Q^T, Q accumulated column-major in q.
    std::vector<double> q(n * n, 0.0);
    for (std::size_t i = 0; i < n; ++i) q[i * n + i] = 1.0;
    auto A = [&](std::size_t r, std::size_t c) -> double& { return a[c * n + r]; };
    std::vector<double> v(n), p(n), w(n);
    for (std::size_t k = 0; k + 2 < n; ++k) {
        double alpha = 0.0;
        for (std::size_t i = k + 1; i < n; ++i) alpha += A(i, k) * A(i, k);
        alpha = std::sqrt(alpha);
        if (alpha == 0.0) continue;
        if (A(k + 1, k) > 0) alpha = -alpha;
        std::fill(v.begin(), v.end(), 0.0);
        for (std::size_t i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (std::size_t i = k + 1; i < n; ++i) vn += v[i] * v[i];
        if (vn == 0.0) continue;
        for (std::size_t i = k + 1; i < n; ++i) v[i] /= std::sqrt(vn);
        // A <- H A H with H = I - 2 v v^T: p = A v, K = v^T p, w = p - K v, A -= 2 (v w^T + w v^T)
        for (std::size_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (std::size_t j = k + 1; j < n; ++j) s += A(i, j) * v[j];
            p[i] = s;
        }
        double K = 0.0;
        for (std::size_t i = k + 1; i < n; ++i) K += v[i] * p[i];
        for (std::size_t i = 0; i < n; ++i) w[i] = p[i] - K * v[i];
        for (std::size_t c = 0; c < n; ++c)
            for (std::size_t r = 0; r < n; ++r) A(r, c) -= 2.0 * (v[r] * w[c] + w[r] * v[c]);
        // Q <- Q H
        for (std::size_t r = 0; r < n; ++r) {
            double s = 0.0;
            for (std::size_t j = k + 1; j < n; ++j) s += q[j * n + r] * v[j];
            for (std::size_t j = k + 1; j < n; ++j) q[j * n + r] -= 2.0 * s * v[j];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (std::size_t i = 0; i < n; ++i) d[i] = A(i, i);
    for (std::size_t i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
    ql_implicit(d, e, n, q.data());
    return sorted(d, &q, n);
}

std::vector<double> hermitian_eigvals(const std::vector<double>& a, std::size_t n) {
    // [[Re, -Im], [Im, Re]] carries every eigenvalue twice.
    const std::size_t N2 = 2 * n;
    std::vector<double> b(N2 * N2, 0.0);
    for (std::size_t c = 0; c < n; ++c)
        for (std::size_t r = 0; r < n; ++r) {
            const double re = a[2 * (c * n + r)], im = a[2 * (c * n + r) + 1];
            b[c * N2 + r] = re;
            b[(c + n) * N2 + (r + n)] = re;
            b[(c + n) * N2 + r] = -im;
            b[c * N2 + (r + n)] = im;
        }
    const TridiagEig t = symmetric_eig(std::move(b), N2);
    std::vector<double> out(n);
    for (std::size_t i = 0; i < n; ++i) out[i] = t.values[2 * i];
    return out;
}

}  // namespace qsa
