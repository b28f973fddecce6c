#include "tridiag.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <complex>
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
// accumulated into the first `rows` rows of the column-major matrix z (leading dimension rows; when
// non-null). Each row is rotated independently, so rows = 1 gives the first row of the full
// eigenvector matrix bit for bit.
void ql_implicit(std::vector<double>& d, std::vector<double>& e, std::size_t n, double* z, std::size_t rows) {
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
                    double* zi = z + ii * rows;
                    double* zj = z + (ii + 1) * rows;
                    for (std::size_t k = 0; k < rows; ++k) {
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
    ql_implicit(d, e, m, vectors ? z.data() : nullptr, m);
    return sorted(d, vectors ? &z : nullptr, m);
}

void tridiag_weights(const double* alpha, const double* beta_next, std::size_t m, std::vector<double>& values,
                     std::vector<double>& weights) {
    values.clear();
    weights.clear();
    if (m == 0) return;
    std::vector<double> d(alpha, alpha + m);
    std::vector<double> e(m, 0.0);
    for (std::size_t i = 0; i + 1 < m; ++i) e[i] = beta_next[i];
    for (std::size_t i = 0; i < m; ++i)
        if (!std::isfinite(d[i]) || !std::isfinite(e[i]))
            throw Error(kConvergence, "tridiagonal eigensolve: non-finite entry at step " + std::to_string(i));
    std::vector<double> z0(m, 0.0);
    z0[0] = 1.0;
    ql_implicit(d, e, m, z0.data(), 1);
    std::vector<std::size_t> order(m);
    std::iota(order.begin(), order.end(), std::size_t{0});
    std::stable_sort(order.begin(), order.end(), [&d](std::size_t a, std::size_t b) { return d[a] < d[b]; });
    values.resize(m);
    weights.resize(m);
    for (std::size_t j = 0; j < m; ++j) {
        values[j] = d[order[j]];
        weights[j] = z0[order[j]] * z0[order[j]];
    }
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
    ql_implicit(d, e, n, q.data(), n);
    return sorted(d, &q, n);
}

std::vector<double> hermitian_eigvals_householder(std::vector<std::complex<double>> a, std::size_t n) {
    using cd = std::complex<double>;
    if (n == 0) return {};
    for (const cd& x : a)
        if (!std::isfinite(x.real()) || !std::isfinite(x.imag()))
            throw Error(kConvergence, "dense eigensolve: non-finite entry");
    // Column-major A; column k below the diagonal is mapped onto a multiple of e_1 by the Hermitian
    // reflector H = I - t u u^H, t = 2 / u^H u, u = x + e^{i arg x_0} ||x|| e_1; the trailing block
    // becomes H A22 H = A22 - u w^H - w u^H with p = t A22 u, K = t u^H p / 2, w = p - K u.
    auto A = [&](std::size_t r, std::size_t c) -> cd& { return a[c * n + r]; };
    std::vector<double> d(n), e(n, 0.0);
    std::vector<cd> u(n), p(n);
    for (std::size_t k = 0; k + 1 < n; ++k) {
        const std::size_t r0 = k + 1;
        double xn2 = 0.0;
        for (std::size_t i = r0; i < n; ++i) xn2 += std::norm(A(i, k));
        const double xn = std::sqrt(xn2);
        e[k] = xn;   // |the subdiagonal entry| after the reflection
        if (k + 2 >= n || xn == 0.0) continue;
        const cd x0 = A(r0, k);
        const double ax0 = std::abs(x0);
        const cd ph = ax0 > 0.0 ? x0 / ax0 : cd(1.0, 0.0);
        for (std::size_t i = r0; i < n; ++i) u[i] = A(i, k);
        u[r0] += ph * xn;
        double un2 = 0.0;
        for (std::size_t i = r0; i < n; ++i) un2 += std::norm(u[i]);
        if (un2 == 0.0) continue;
        const double t = 2.0 / un2;
        for (std::size_t i = r0; i < n; ++i) p[i] = 0.0;
        for (std::size_t c = r0; c < n; ++c) {
            const cd uc = u[c];
            const cd* col = &a[c * n];
            for (std::size_t r = r0; r < n; ++r) p[r] += col[r] * uc;
        }
        cd up(0.0, 0.0);
        for (std::size_t i = r0; i < n; ++i) {
            p[i] *= t;
            up += std::conj(u[i]) * p[i];
        }
        const double K = 0.5 * t * up.real();
        for (std::size_t i = r0; i < n; ++i) p[i] -= K * u[i];
        for (std::size_t c = r0; c < n; ++c) {
            const cd wc = std::conj(p[c]), uc = std::conj(u[c]);
            cd* col = &a[c * n];
            for (std::size_t r = r0; r < n; ++r) col[r] -= u[r] * wc + p[r] * uc;
        }
    }
    for (std::size_t k = 0; k < n; ++k) d[k] = A(k, k).real();
    std::vector<double> beta(n + 1, 0.0);
    for (std::size_t k = 0; k + 1 < n; ++k) beta[k + 1] = e[k];
    return tridiag_eig(d, beta, n, false).values;
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
