// CPU ORACLE (test infrastructure only; never linked into the product).
//
// Reference-element arithmetic restated from the reference's source text:
// equispaced Lagrange basis, tensor products (dimension 0 slowest),
// Gauss-Legendre quadrature in the reference's point order, and the small
// dense determinant/adjugate helpers that the absent numericaltoolbox
// submodule supplied (Numtool/fixed_size_tensor.hpp call sites in
// finite_element.hpp:235-236, 354-357).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace oracle {

constexpr int MAXP = 6;

// UniformLagrangeInterpolation<T, Pn> (basis/lagrange_1d.hpp:19-212):
// nodes -1 + cumulative 2/p steps, barycentric weights, evaluation that skips
// the node nearest to xi, P1 hard-coded.
struct Lagrange1D {
    int p = 1;
    double x[MAXP + 1]{};
    double w[MAXP + 1]{};

    explicit Lagrange1D(int p_) : p(p_) {
        if (p == 0) {
            x[0] = 0.0;
        } else {
            const double dx = 2.0 / p;
            x[0] = -1.0;
            for (int j = 1; j <= p; ++j) x[j] = x[j - 1] + dx;  // lagrange_1d.hpp:40-45
        }
        for (int j = 0; j <= p; ++j) {
            w[j] = 1.0;
            for (int k = 0; k <= p; ++k)
                if (k != j) w[j] *= (x[j] - x[k]);
            w[j] = 1.0 / w[j];
        }
    }

    // lagrange_1d.hpp:77-117
    void eval(double xi, double* N) const {
        if (p == 0) { N[0] = 1.0; return; }
        if (p == 1) { N[0] = 0.5 * (1 - xi); N[1] = 1.0 - N[0]; return; }
        double lskip = 1;
        int k;
        for (k = 0; k < p; ++k) {
            if (xi >= (x[k] + x[k + 1]) / 2) lskip *= (xi - x[k]);
            else break;
        }
        for (int i = k + 1; i < p + 1; ++i) lskip *= (xi - x[i]);
        const double lprod = lskip * (xi - x[k]);
        int j;
        for (j = 0; j < k; ++j) N[j] = lprod * w[j] / (xi - x[j]);
        N[k] = lskip * w[k];
        for (++j; j < p + 1; ++j) N[j] = lprod * w[j] / (xi - x[j]);
    }

    // lagrange_1d.hpp:126-193
    void deriv(double xi, double* N, double* dN) const {
        if (p == 0) { N[0] = 1.0; dN[0] = 0.0; return; }
        if (p == 1) {
            N[0] = 0.5 * (1 - xi); N[1] = 1.0 - N[0];
            dN[0] = -0.5; dN[1] = 0.5;
            return;
        }
        double lskip = 1;
        int k;
        for (k = 0; k < p; ++k) {
            if (xi >= (x[k] + x[k + 1]) / 2) lskip *= (xi - x[k]);
            else break;
        }
        for (int i = k + 1; i < p + 1; ++i) lskip *= (xi - x[i]);
        const double lprod = lskip * (xi - x[k]);
        double s = 0.0;
        int j;
        for (j = 0; j < k; ++j) {
            const double inv = 1.0 / (xi - x[j]);
            s += inv;
            N[j] = lprod * inv * w[j];
        }
        N[k] = lskip * w[k];
        for (++j; j < p + 1; ++j) {
            const double inv = 1.0 / (xi - x[j]);
            s += inv;
            N[j] = lprod * inv * w[j];
        }
        const double lprime = lprod * s + lskip;
        for (j = 0; j < k; ++j) dN[j] = (lprime * w[j] - N[j]) / (xi - x[j]);
        dN[k] = s * N[k];
        for (++j; j < p + 1; ++j) dN[j] = (lprime * w[j] - N[j]) / (xi - x[j]);
    }

    // Second derivative (the reference calls Numtool's dNlagrange1d<T,Pn>(j,2,xi),
    // lagrange_1d.hpp:209, which is absent): l_j'' = w_j * sum over ordered pairs
    // a != b (both != j) of prod_{l != j,a,b} (xi - x_l).  Pinned by the closed
    // forms of TestHypercubeTransform.cpp:705-776 (P2).
    double second(int j, double xi) const {
        if (p < 2) return 0.0;
        double s = 0.0;
        for (int a = 0; a <= p; ++a) {
            if (a == j) continue;
            for (int b = 0; b <= p; ++b) {
                if (b == j || b == a) continue;
                double prod = 1.0;
                for (int l = 0; l <= p; ++l)
                    if (l != j && l != a && l != b) prod *= (xi - x[l]);
                s += prod;
            }
        }
        return s * w[j];
    }

    void d2(double xi, double* N, double* dN, double* d2N) const {
        deriv(xi, N, dN);
        for (int j = 0; j <= p; ++j) d2N[j] = second(j, xi);
    }
};

// Legendre polynomial and derivative by the three-term recurrence (the
// reference calls MATH::POLYNOMIAL::legendre1d / dlegendre1d from the absent
// numericaltoolbox, quadrules_1d.hpp:50-51).
inline void legendre(int n, double x, double& P, double& dP) {
    if (n == 0) { P = 1.0; dP = 0.0; return; }
    double p0 = 1.0, p1 = x;
    for (int k = 1; k < n; ++k) {
        const double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
    }
    P = p1;
    dP = n * (x * p1 - p0) / (x * x - 1.0);
}

// GaussLegendreQuadrature<T, IDX, npoin> (quadrature/quadrules_1d.hpp:20-82):
// points ordered (+x1, -x1, +x2, -x2, ..., 0), Newton from
// cos(pi (i - 1/4) / (n + 1/2)), centre weight 2 - sum of the others.
struct GaussLegendre1D {
    int n = 1;
    double x[MAXP + 2]{};
    double w[MAXP + 2]{};
    explicit GaussLegendre1D(int n_) : n(n_) {
        const int m = n / 2;
        double weightcenter = 2.0;
        for (int i = 1; i <= m; ++i) {
            double xi = std::cos(M_PI * (i - 0.25) / (n + 0.5));
            double dleg = 0.0;
            for (int it = 0; it < 50; ++it) {
                double leg;
                legendre(n, xi, leg, dleg);
                const double dxi = leg / dleg;
                if (std::abs(dxi) < 1e-16) { xi -= dxi; break; }
                xi -= dxi;
            }
            const double weight = 2.0 / ((1 - xi * xi) * dleg * dleg);
            if (n % 2 != 0) weightcenter -= 2 * weight;
            x[2 * (i - 1)] = xi;
            w[2 * (i - 1)] = weight;
            x[2 * (i - 1) + 1] = -xi;
            w[2 * (i - 1) + 1] = weight;
        }
        if (n % 2 != 0) { x[n - 1] = 0.0; w[n - 1] = weightcenter; }
    }
};

// Index helpers for tensor products with dimension 0 slowest
// (QTypeProduct::ijk_poin / convert_ijk, tensor_product.hpp:336-380).
template <int D>
inline std::array<int, D> tensor_ijk(int idx, int n) {
    std::array<int, D> ijk{};
    for (int d = D - 1; d >= 0; --d) { ijk[d] = idx % n; idx /= n; }
    return ijk;
}
inline int ipow(int b, int e) { int r = 1; for (int i = 0; i < e; ++i) r *= b; return r; }

// HypercubeLagrangeBasis (basis/lagrange.hpp:88-131) built on QTypeProduct:
// values, reference gradients [nb][D] and Hessians [nb][D][D].
template <int D>
struct TensorBasis {
    Lagrange1D l1;
    int n, nb;
    std::vector<std::array<int, D>> ijk;
    explicit TensorBasis(int p) : l1(p), n(p + 1), nb(ipow(p + 1, D)) {
        for (int i = 0; i < nb; ++i) ijk.push_back(tensor_ijk<D>(i, n));
    }
    void eval(const double* xi, double* B, double* G, double* H) const {
        double N[D][MAXP + 1], dN[D][MAXP + 1], d2N[D][MAXP + 1];
        for (int d = 0; d < D; ++d) l1.d2(xi[d], N[d], dN[d], d2N[d]);
        for (int i = 0; i < nb; ++i) {
            const auto& c = ijk[i];
            double v = 1.0;
            for (int d = 0; d < D; ++d) v *= N[d][c[d]];
            if (B) B[i] = v;
            if (G) {
                for (int k = 0; k < D; ++k) {
                    double g = 1.0;
                    for (int d = 0; d < D; ++d) g *= (d == k) ? dN[d][c[d]] : N[d][c[d]];
                    G[i * D + k] = g;
                }
            }
            if (H) {
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D; ++b) {
                        double h = 1.0;
                        for (int d = 0; d < D; ++d) {
                            if (a == b) h *= (d == a) ? d2N[d][c[d]] : N[d][c[d]];
                            else h *= (d == a || d == b) ? dN[d][c[d]] : N[d][c[d]];
                        }
                        H[(i * D + a) * D + b] = h;
                    }
            }
        }
    }
};

// HypercubeGaussLegendre<T, IDX, ndim, npoin1d> (HypercubeGaussLegendre.hpp:17-67).
template <int D>
struct TensorQuadrature {
    int npts;
    std::vector<std::array<double, (D > 0 ? D : 1)>> pts;
    std::vector<double> w;
    explicit TensorQuadrature(int n1d) {
        GaussLegendre1D g(n1d);
        npts = ipow(n1d, D);
        pts.resize(npts);
        w.assign(npts, 1.0);
        if constexpr (D == 0) {
            pts[0][0] = 0.0;
            w[0] = 1.0;
        } else {
            for (int q = 0; q < npts; ++q) {
                auto ijk = tensor_ijk<D>(q, n1d);
                for (int d = 0; d < D; ++d) {
                    pts[q][d] = g.x[ijk[d]];
                    if (d == 0) w[q] = g.w[ijk[d]];
                    else w[q] *= g.w[ijk[d]];
                }
            }
        }
    }
};

// ---- dense helpers for D <= 3 -----------------------------------------
template <int D>
inline double determinant(const double (&J)[D][D]) {
    if constexpr (D == 1) return J[0][0];
    else if constexpr (D == 2) return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
        return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
             - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
             + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// adjugate: J * adj(J) = det(J) I
template <int D>
inline void adjugate(const double (&J)[D][D], double (&A)[D][D]) {
    if constexpr (D == 1) {
        A[0][0] = 1.0;
    } else if constexpr (D == 2) {
        A[0][0] = J[1][1]; A[0][1] = -J[0][1];
        A[1][0] = -J[1][0]; A[1][1] = J[0][0];
    } else {
        A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
        A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
        A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    }
}

// Dense LU with partial pivoting and solve (the reference uses Numtool
// decompose_lu / sub_lu, ssp_rk3.hpp:145-157; that pivoting is not
// available to pin, results agree to rounding).
struct DenseLU {
    int n = 0;
    std::vector<double> a;
    std::vector<int> piv;
    bool singular = false;
    void factor(const std::vector<double>& m, int n_) {
        n = n_;
        a = m;
        piv.resize(n);
        singular = false;
        for (int i = 0; i < n; ++i) piv[i] = i;
        for (int k = 0; k < n; ++k) {
            int pk = k;
            double best = std::abs(a[k * n + k]);
            for (int i = k + 1; i < n; ++i)
                if (std::abs(a[i * n + k]) > best) { best = std::abs(a[i * n + k]); pk = i; }
            if (best == 0.0) { singular = true; return; }
            if (pk != k) {
                for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[pk * n + j]);
                std::swap(piv[k], piv[pk]);
            }
            for (int i = k + 1; i < n; ++i) {
                const double f = a[i * n + k] / a[k * n + k];
                a[i * n + k] = f;
                for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
            }
        }
    }
    void solve(const double* b, double* x) const {
        std::vector<double> y(n);
        for (int i = 0; i < n; ++i) {
            double s = b[piv[i]];
            for (int j = 0; j < i; ++j) s -= a[i * n + j] * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
            x[i] = s / a[i * n + i];
        }
    }
};

}  // namespace oracle
