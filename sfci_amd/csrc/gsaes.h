// Dense kernels of the Krylov drivers, host code, header-only.
//  gsaes  generalized symmetric eigenproblem K x = lambda S x for the Nl x Nl projected
//         Lanczos matrices (the reference uses Eigen's GeneralizedSelfAdjointEigenSolver,
//         src/sparse_solver.cpp:287-288, 450-454): Cholesky S = L L^T, cyclic Jacobi on
//         C = L^-1 K L^-T, x = L^-T y.  Eigenvalues ascending; eigenvectors S-normalized
//         with the first nonzero component made positive.
//  lstsq  least-squares solution of the (n+1) x n system of the Krylov linear solver (the
//         reference uses Eigen's colPivHouseholderQr().solve, src/sparse_solver.cpp:672):
//         Householder QR with column pivoting (Businger-Golub), rank decided as Eigen
//         does (|R_kk| > eps * min(m, n) * |R_00|), free variables set to zero.
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace sfci_dense {

// Generalized symmetric eigenproblem K x = lambda S x (n x n, row-major).
// Returns 0, or -1 when S is not positive definite.
inline int gsaes(int n, const std::vector<double>& K, const std::vector<double>& S,
          std::vector<double>& evals, std::vector<double>& evecs) {
    std::vector<double> L(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = S[i * n + j];
            for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
            if (i == j) {
                if (!(s > 0.0)) return -1;
                L[i * n + i] = std::sqrt(s);
            } else {
                L[i * n + j] = s / L[j * n + j];
            }
        }
    // C = L^-1 K L^-T : first Y = L^-1 K (forward substitution per column)
    std::vector<double> Y(n * n), C(n * n);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) {
            double s = K[i * n + c];
            for (int k = 0; k < i; ++k) s -= L[i * n + k] * Y[k * n + c];
            Y[i * n + c] = s / L[i * n + i];
        }
    // C = Y L^-T  <=>  C^T = L^-1 Y^T
    for (int r = 0; r < n; ++r)
        for (int i = 0; i < n; ++i) {
            double s = Y[r * n + i];
            for (int k = 0; k < i; ++k) s -= L[i * n + k] * C[r * n + k];
            C[r * n + i] = s / L[i * n + i];
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            double a = 0.5 * (C[i * n + j] + C[j * n + i]);
            C[i * n + j] = C[j * n + i] = a;
        }
    // cyclic Jacobi
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += C[i * n + j] * C[i * n + j];
                if (i != j) off += C[i * n + j] * C[i * n + j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = C[p * n + q];
                if (apq == 0.0) continue;
                double app = C[p * n + p], aqq = C[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double ckp = C[k * n + p], ckq = C[k * n + q];
                    C[k * n + p] = c * ckp - s * ckq;
                    C[k * n + q] = s * ckp + c * ckq;
                }
                for (int k = 0; k < n; ++k) {
                    double cpk = C[p * n + k], cqk = C[q * n + k];
                    C[p * n + k] = c * cpk - s * cqk;
                    C[q * n + k] = s * cpk + c * cqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return C[a * n + a] < C[b * n + b]; });
    evals.assign(n, 0.0);
    evecs.assign(n * n, 0.0);   // column j = eigenvector j
    for (int jj = 0; jj < n; ++jj) {
        int j = order[jj];
        evals[jj] = C[j * n + j];
        // x = L^-T y
        std::vector<double> x(n);
        for (int i = n - 1; i >= 0; --i) {
            double s = V[i * n + j];
            for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * x[k];
            x[i] = s / L[i * n + i];
        }
        double sg = 1.0;
        for (int i = 0; i < n; ++i)
            if (x[i] != 0.0) { sg = x[i] > 0 ? 1.0 : -1.0; break; }
        for (int i = 0; i < n; ++i) evecs[i * n + jj] = sg * x[i];
    }
    return 0;
}

// min ||A x - b||_2 for A (m x n, row-major).  Returns the numerical rank.
inline int lstsq(int m, int n, const std::vector<double>& A, const std::vector<double>& b, std::vector<double>& x) {
    std::vector<double> R(A), y(b);
    std::vector<int> perm(n);
    for (int j = 0; j < n; ++j) perm[j] = j;
    const int d = std::min(m, n);
    double r00 = 0.0;
    int rank = 0;
    for (int k = 0; k < d; ++k) {
        // pivot: the remaining column of largest norm (recomputed, no downdating)
        int p = k;
        double best = -1.0;
        for (int j = k; j < n; ++j) {
            double s = 0.0;
            for (int i = k; i < m; ++i) s += R[i * n + j] * R[i * n + j];
            if (s > best) { best = s; p = j; }
        }
        if (p != k) {
            for (int i = 0; i < m; ++i) std::swap(R[i * n + k], R[i * n + p]);
            std::swap(perm[k], perm[p]);
        }
        double nrm = std::sqrt(best);
        if (k == 0) r00 = nrm;
        if (!(nrm > std::numeric_limits<double>::epsilon() * d * r00) || nrm == 0.0) break;
        ++rank;
        double alpha = R[k * n + k] > 0 ? -nrm : nrm;
        std::vector<double> v(m - k);
        for (int i = k; i < m; ++i) v[i - k] = R[i * n + k];
        v[0] -= alpha;
        double vv = 0.0;
        for (double t : v) vv += t * t;
        if (vv > 0.0) {
            for (int j = k; j < n; ++j) {
                double s = 0.0;
                for (int i = k; i < m; ++i) s += v[i - k] * R[i * n + j];
                s = 2.0 * s / vv;
                for (int i = k; i < m; ++i) R[i * n + j] -= s * v[i - k];
            }
            double s = 0.0;
            for (int i = k; i < m; ++i) s += v[i - k] * y[i];
            s = 2.0 * s / vv;
            for (int i = k; i < m; ++i) y[i] -= s * v[i - k];
        }
    }
    std::vector<double> z(n, 0.0);
    for (int i = rank - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < rank; ++j) s -= R[i * n + j] * z[j];
        z[i] = s / R[i * n + i];
    }
    x.assign(n, 0.0);
    for (int j = 0; j < n; ++j) x[perm[j]] = z[j];
    return rank;
}

}  // namespace sfci_dense
