// Host construction of the reference-element tables shared by the 2D and 1D
// device contexts (nodal_basis.hpp:13-65).
#include <algorithm>
#include <cmath>
#include <vector>

#include "dgh_host.hpp"
#include "dgh_internal.hpp"

namespace dgh {

// ------------------------------------------------------------- host basis
// Gauss-Legendre nodes/weights by Newton iteration on P_n, ascending, exactly
// symmetric, middle node 0 for odd n (nodal_basis.hpp:13-25 conventions).
void gauss_rule(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double pi = std::acos(-1.0);
    auto legendre = [n](double z, double &pn, double &dpn) {
        double p0 = 1.0, p1 = z;
        for (int l = 2; l <= n; ++l) {
            const double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) { p0 = 1.0; p1 = z; }
        pn = p1;
        dpn = n * (z * p1 - p0) / (z * z - 1.0);
    };
    for (int k = 1; k <= (n + 1) / 2; ++k) {
        double z = std::cos(pi * (k - 0.25) / (n + 0.5)), pn, dpn;
        for (int it = 0; it < 100; ++it) {
            legendre(z, pn, dpn);
            const double dz = pn / dpn;
            z -= dz;
            if (std::abs(dz) <= 1e-16 * std::max(1.0, std::abs(z)) && it > 2) break;
        }
        legendre(z, pn, dpn);
        x[k - 1] = -z;
        x[n - k] = z;
        w[k - 1] = w[n - k] = 2.0 / ((1.0 - z * z) * dpn * dpn);
    }
    if (n % 2) x[n / 2] = 0;
}

std::vector<double> lagrange_at(const std::vector<double> &x, double z) {
    const int n = static_cast<int>(x.size());
    std::vector<double> v(n, 1.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) v[i] *= (z - x[j]) / (x[i] - x[j]);
    return v;
}

// nodal_basis.hpp:34-65: the nodal rule, plus the Q-point volume rule when
// over-integrating (Q > n).
Basis make_basis(int n, double hx, double hy, int Q) {
    Basis b{};
    std::vector<double> x, w;
    gauss_rule(n, x, w);
    std::vector<double> bary(n, 1.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) bary[i] /= x[i] - x[j];
    std::vector<double> D(n * n, 0.0);  // D(i, j) at [i + n*j]
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j)
            if (i != j) D[i + n * j] = bary[j] / (bary[i] * (x[i] - x[j]));
        for (int j = 0; j < n; ++j) s += D[i + n * j];
        D[i + n * i] = -s;
    }
    for (int q = 0; q < n; ++q)
        for (int i = 0; i < n; ++i) b.wd[q + n * i] = w[q] * D[q + n * i];
    for (int i = 0; i < n; ++i) { b.w[i] = w[i]; b.x[i] = x[i]; }
    const auto l = lagrange_at(x, -1.0), r = lagrange_at(x, 1.0);
    for (int i = 0; i < n; ++i) { b.left[i] = l[i]; b.right[i] = r[i]; }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            b.mass[i + n * j] = .25 * hx * hy * w[i] * w[j];
            b.inv_mass[i + n * j] = 1.0 / b.mass[i + n * j];
        }
    const int eq = std::max(20, 2 * n + 3);
    std::vector<double> ex, ew;
    gauss_rule(eq, ex, ew);
    for (int q = 0; q < eq; ++q) {
        b.ex[q] = ex[q];
        b.ew[q] = ew[q];
        const auto e = lagrange_at(x, ex[q]);
        for (int i = 0; i < n; ++i) b.ev[q + eq * i] = e[i];
    }
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < eq; ++q) b.proj[i + n * q] = ((1.0 / w[i]) * b.ev[q + eq * i]) * ew[q];
    if (Q > n) {
        std::vector<double> qx, qw;
        gauss_rule(Q, qx, qw);
        for (int q = 0; q < Q; ++q) {
            b.qw[q] = qw[q];
            const auto e = lagrange_at(x, qx[q]);
            for (int k = 0; k < n; ++k) b.vq[q + Q * k] = e[k];
        }
        for (int q = 0; q < Q; ++q)
            for (int j = 0; j < n; ++j) {
                double s = 0;
                for (int k = 0; k < n; ++k) s += (qw[q] * b.vq[q + Q * k]) * D[k + n * j];
                b.dq[q + Q * j] = s;
            }
    }
    return b;
}

}  // namespace dgh
