// model_math.h -- host-side FP64 model mathematics for libphylo_hip:
// symmetric eigen-decomposition (Householder tridiagonalisation + implicit QL),
// P(t) by eigen-system, by scaling-and-squaring Pade, and by the reference's
// (I + Q t / 2^20)^(2^20) squaring (marth.cpp:129-140, sub_model.cpp:115-127).
//
// Published methods followed: the symmetric eigen-solver is the classical
// Householder tridiagonalisation with accumulated transforms followed by the
// QL algorithm with implicit shifts, as in EISPACK's TRED2 and TQL2 (Martin,
// Reinsch and Wilkinson, Numer. Math. 11 (1968) 181-195; Bowdler, Martin,
// Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306).  The Pade form is
// the degree-13 scaling-and-squaring method of Higham, SIAM J. Matrix Anal.
// Appl. 26 (2005) 1179-1193.
#pragma once

#include <cmath>
#include <cstring>
#include <vector>

namespace phylo_math {

// Householder reduction of the symmetric n x n matrix a (row-major, in place)
// to tridiagonal form; a becomes the accumulated orthogonal transform, d the
// diagonal and e the sub-diagonal (e[0] = 0).
inline void householder_tridiag(int n, std::vector<double>& a, std::vector<double>& d, std::vector<double>& e) {
    auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    for (int i = n - 1; i > 0; --i) {
        int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            double scale = 0.0;
            for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
            if (scale == 0.0) {
                e[i] = A(i, l);
            } else {
                for (int k = 0; k <= l; ++k) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
                double f = A(i, l);
                double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                A(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    A(j, i) = A(i, j) / h;
                    double gg = 0.0;
                    for (int k = 0; k <= j; ++k) gg += A(j, k) * A(i, k);
                    for (int k = j + 1; k <= l; ++k) gg += A(k, j) * A(i, k);
                    e[j] = gg / h;
                    f += e[j] * A(i, j);
                }
                double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    double ff = A(i, j);
                    double gg = e[j] - hh * ff;
                    e[j] = gg;
                    for (int k = 0; k <= j; ++k) A(j, k) -= ff * e[k] + gg * A(i, k);
                }
            }
        } else {
            e[i] = A(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
                for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
            }
        }
        d[i] = A(i, i);
        A(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
    }
}

// Implicit-shift QL on the tridiagonal (d, e); z accumulates eigenvectors
// (columns).  Returns false if it fails to converge.
inline bool tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
    auto Z = [&](int i, int j) -> double& { return z[(size_t)i * n + j]; };
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * f;
                        Z(k, i) = c * Z(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return true;
}

// Eigen-system of a reversible Q with stationary pi: Q = V diag(lam) Vinv with
// V = D^-1/2 U, Vinv = U^T D^1/2, U the eigenvectors of D^1/2 Q D^-1/2.
inline bool reversible_eigen(int S, const double* Q, const double* pi, std::vector<double>& V,
                             std::vector<double>& lam, std::vector<double>& Vinv) {
    std::vector<double> B((size_t)S * S), d, e, sq(S);
    for (int i = 0; i < S; ++i) {
        if (!(pi[i] > 0.0)) return false;
        sq[i] = std::sqrt(pi[i]);
    }
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) B[(size_t)i * S + j] = sq[i] * Q[(size_t)i * S + j] / sq[j];
    for (int i = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j) {
            double a = 0.5 * (B[(size_t)i * S + j] + B[(size_t)j * S + i]);
            B[(size_t)i * S + j] = B[(size_t)j * S + i] = a;
        }
    householder_tridiag(S, B, d, e);
    if (!tridiag_ql(S, d, e, B)) return false;
    V.assign((size_t)S * S, 0.0);
    Vinv.assign((size_t)S * S, 0.0);
    lam = d;
    for (int i = 0; i < S; ++i)
        for (int k = 0; k < S; ++k) {
            V[(size_t)i * S + k] = B[(size_t)i * S + k] / sq[i];
            Vinv[(size_t)k * S + i] = B[(size_t)i * S + k] * sq[i];
        }
    return true;
}

// Detailed balance check pi_i Q_ij == pi_j Q_ji (relative to the largest rate).
inline bool is_reversible(int S, const double* Q, const double* pi) {
    double mx = 0.0;
    for (int i = 0; i < S * S; ++i) mx = std::fmax(mx, std::fabs(Q[i]));
    for (int i = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j)
            if (std::fabs(pi[i] * Q[(size_t)i * S + j] - pi[j] * Q[(size_t)j * S + i]) > 1e-12 * (mx + 1e-300))
                return false;
    return true;
}

inline void matmul(int n, const double* A, const double* B, double* C) {
#pragma clang fp contract(off)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * B[(size_t)k * n + j];
            C[(size_t)i * n + j] = s;
        }
}

// P = exp(Q t) by scaling and squaring with the [13/13] Pade approximant
// (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193).
inline void expm_pade13(int n, const double* Q, double t, double* P) {
    static const double b[14] = {64764752532480000., 32382376266240000., 7771770303897600.,
                                 1187353796428800., 129060195264000., 10559470521600., 670442572800.,
                                 33522128640., 1323241920., 40840800., 960960., 16380., 182., 1.};
    const size_t nn = (size_t)n * n;
    std::vector<double> buf(nn * 10);
    double *A = buf.data(), *A2 = A + nn, *A4 = A2 + nn, *A6 = A4 + nn, *U = A6 + nn, *V = U + nn,
           *T = V + nn, *W = T + nn, *N = W + nn, *D = N + nn;
    double norm = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += std::fabs(Q[(size_t)i * n + j] * t);
        norm = std::fmax(norm, s);
    }
    int sq = 0;
    if (norm > 5.371920351148152) sq = (int)std::ceil(std::log2(norm / 5.371920351148152));
    const double sc = std::ldexp(1.0, -sq);
    for (size_t i = 0; i < nn; ++i) A[i] = Q[i] * t * sc;
    matmul(n, A, A, A2);
    matmul(n, A2, A2, A4);
    matmul(n, A4, A2, A6);
    for (size_t i = 0; i < nn; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    matmul(n, A6, T, W);
    for (size_t i = 0; i < nn; ++i) W[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
    for (int i = 0; i < n; ++i) W[(size_t)i * n + i] += b[1];
    matmul(n, A, W, U);
    for (size_t i = 0; i < nn; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    matmul(n, A6, T, V);
    for (size_t i = 0; i < nn; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] += b[0];
    for (size_t i = 0; i < nn; ++i) { N[i] = V[i] + U[i]; D[i] = V[i] - U[i]; }
    // solve D X = N by Gauss-Jordan with partial pivoting
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(D[(size_t)r * n + c]) > std::fabs(D[(size_t)piv * n + c])) piv = r;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(D[(size_t)c * n + k], D[(size_t)piv * n + k]);
                std::swap(N[(size_t)c * n + k], N[(size_t)piv * n + k]);
            }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = D[(size_t)r * n + c] / D[(size_t)c * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                D[(size_t)r * n + k] -= f * D[(size_t)c * n + k];
                N[(size_t)r * n + k] -= f * N[(size_t)c * n + k];
            }
        }
    }
    for (int r = 0; r < n; ++r) {
        double dd = D[(size_t)r * n + r];
        for (int k = 0; k < n; ++k) N[(size_t)r * n + k] /= dd;
    }
    for (int s = 0; s < sq; ++s) {
        matmul(n, N, N, T);
        std::memcpy(N, T, sizeof(double) * nn);
    }
    std::memcpy(P, N, sizeof(double) * nn);
}

// The reference's P(t): I + Q t / 2^20, squared 20 times, products summed in
// ascending k from 0 (marth.cpp:50-61, :106-111, :129-140).
inline void p_refcompat(int n, const double* Q, double t, double* P) {
#pragma clang fp contract(off)
    const size_t nn = (size_t)n * n;
    std::vector<double> T(nn);
    const double scale = 1048576.0;
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y) P[(size_t)x * n + y] = (x == y ? 1.0 : 0.0) + (Q[(size_t)x * n + y] * t) / scale;
    for (int i = 0; i < 20; ++i) {
        matmul(n, P, P, T.data());
        std::memcpy(P, T.data(), sizeof(double) * nn);
    }
}

}  // namespace phylo_math
