// C++/OpenMP restatement of the reference's dense-mesh shift / injection current k-loop
// (TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a second, independent implementation that
// tests compare with the NumPy oracle oracle/kloop.py; nothing in the product path loads it).
//
// Restated reference functions (paths relative to the reference root), the same as oracle/kloop.py:
//   * Fourier sums (Convention II, displacement = Cartesian R): HamiltonianMatrixElements.jl:28-58,92-177,
//     PositionMatrixElements.jl:8-84,239-409.  Evaluated on a uniform mesh plane by plane, as the reference's
//     Mixed-FFT backend does for mesh runs: X(k) = sum_R1 e1 sum_R2 e2 sum_R3 e3 X(R) with the R3 sum once
//     per k3 plane and the R2 sum once per k2 line (the same sums, reordered).
//   * Eigensystem of the upper triangle (DenseSpectra.jl:14-16, LAPACK uplo 'U'): Householder reduction of
//     the Hermitian matrix to a real symmetric tridiagonal one (the unblocked zhetd2 steps), implicit QL
//     with Wilkinson shifts on the tridiagonal matrix, back-transformation; ascending order (ties by
//     index); canonical phase per column, largest |v_i| real and non-negative, first on ties
//     (HamiltonianMatrixElements.jl:6-25).
//   * U^+ X U of every channel (HamiltonianMatrixElements.jl:61-71); D, Berry connection
//     (PositionVertexRegularization.jl:17-29, PositionMatrixElements.jl:444-489).
//   * Occupations and screening (CalculationUtils.jl:21-45, TransitionScreening.jl:32-94; band window -1).
//   * Generalized derivative (ConventionalShiftCurrent.jl:191-229), shift-current kernel (:710-729),
//     injection kernel (ConventionalInjectionCurrent.jl:23-97), frequency contraction (:774-884).
// Output: the raw sums over the points of the range of sum_pairs (f_n - f_m) delta(w) Q[a][b][c] (no
// prefactor, no 1/N_k: the caller applies them as oracle/kloop.py does).
// A point without an active pair contributes nothing and is skipped after the screening.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

typedef std::complex<double> cd;
constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltzmannEvK = 8.617333262e-5;
constexpr int kChannels = 22;   // H, dH[3], ddH[6: xx yy zz yz xz xy], A[3], dA[9: 3a + b]

inline int sym_index(int a, int b) {
    if (a == b) return a;
    const int s = a + b;
    return s == 3 ? 3 : (s == 2 ? 4 : 5);
}

// ------------------------------------------------------------------------------------------------
// Hermitian eigensystem: A (full, row-major, n x n) is overwritten; E ascending, V columns.
// ------------------------------------------------------------------------------------------------
void hermitian_eigen(int n, std::vector<cd> &A, std::vector<double> &E, std::vector<cd> &V) {
    std::vector<double> d(n), e(n, 0.0);
    std::vector<cd> tau(n, cd(0.0, 0.0)), w(n);
    auto a = [&](int i, int j) -> cd & { return A[(size_t)i * n + j]; };
    // Householder reduction (lower triangle): for column i, H_i = I - tau v v^H with v = (1, x / (alpha -
    // beta)) annihilates A(i+2:n, i); A(i+1:n, i+1:n) <- H_i^H A H_i as a rank-2 update.
    for (int i = 0; i + 1 < n; ++i) {
        const int m = n - i - 1;   // length of v
        cd alpha = a(i + 1, i);
        double xnorm2 = 0.0;
        for (int k = i + 2; k < n; ++k) xnorm2 += std::norm(a(k, i));
        cd t(0.0, 0.0);
        if (xnorm2 == 0.0 && alpha.imag() == 0.0) {
            e[i] = alpha.real();
        } else {
            const double norm = std::sqrt(std::norm(alpha) + xnorm2);
            const double beta = alpha.real() >= 0.0 ? -norm : norm;
            t = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
            const cd scale = 1.0 / (alpha - beta);
            for (int k = i + 2; k < n; ++k) a(k, i) *= scale;
            e[i] = beta;
            a(i + 1, i) = 1.0;
            // w = t A22 v, then w -= (t/2)(w^H v) v, A22 -= v w^H + w v^H
            for (int r = 0; r < m; ++r) {
                cd s(0.0, 0.0);
                for (int c = 0; c < m; ++c) s += a(i + 1 + r, i + 1 + c) * a(i + 1 + c, i);
                w[r] = t * s;
            }
            cd dot(0.0, 0.0);
            for (int r = 0; r < m; ++r) dot += std::conj(w[r]) * a(i + 1 + r, i);
            const cd al = -0.5 * t * dot;
            for (int r = 0; r < m; ++r) w[r] += al * a(i + 1 + r, i);
            for (int r = 0; r < m; ++r)
                for (int c = 0; c < m; ++c)
                    a(i + 1 + r, i + 1 + c) -= a(i + 1 + r, i) * std::conj(w[c]) + w[r] * std::conj(a(i + 1 + c, i));
        }
        tau[i] = t;
        d[i] = a(i, i).real();
    }
    d[n - 1] = a(n - 1, n - 1).real();
    // implicit QL on (d, e): e[i] couples i and i + 1; Z accumulates the plane rotations (real)
    std::vector<double> Z((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 60; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m == l) break;
            // Wilkinson shift from the leading 2 x 2 block at l
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool underflow = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                const double rr = (d[i] - g) * s + 2.0 * c * b;
                p = s * rr;
                d[i + 1] = g + p;
                g = c * rr - b;
                for (int k = 0; k < n; ++k) {
                    double &z0 = Z[(size_t)k * n + i], &z1 = Z[(size_t)k * n + i + 1];
                    const double zf = z1;
                    z1 = s * z0 + c * zf;
                    z0 = c * z0 - s * zf;
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // eigenvectors of A: Q Z with Q = H_0 H_1 ... H_{n-2}; apply H_i (acting on rows i+1..n-1) from the last
    std::vector<cd> Q((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Q[(size_t)i * n + j] = cd(Z[(size_t)i * n + j], 0.0);
    for (int i = n - 2; i >= 0; --i) {
        if (tau[i] == cd(0.0, 0.0)) continue;
        const int m = n - i - 1;
        for (int j = 0; j < n; ++j) {
            cd s(0.0, 0.0);
            for (int r = 0; r < m; ++r) s += std::conj(a(i + 1 + r, i)) * Q[(size_t)(i + 1 + r) * n + j];
            s *= tau[i];
            for (int r = 0; r < m; ++r) Q[(size_t)(i + 1 + r) * n + j] -= a(i + 1 + r, i) * s;
        }
    }
    // ascending order, ties by index
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    E.resize(n);
    V.assign((size_t)n * n, cd(0.0, 0.0));
    for (int j = 0; j < n; ++j) {
        const int src = order[j];
        E[j] = d[src];
        int piv = 0;
        double pm = std::abs(Q[src]);
        for (int i = 1; i < n; ++i) {
            const double mag = std::abs(Q[(size_t)i * n + src]);
            if (mag > pm) {
                pm = mag;
                piv = i;
            }
        }
        const cd ph = pm > 0.0 ? std::conj(Q[(size_t)piv * n + src]) / pm : cd(1.0, 0.0);
        for (int i = 0; i < n; ++i) V[(size_t)i * n + j] = Q[(size_t)i * n + src] * ph;
    }
}

// U^+ X U (row-major, n x n), planar real / imaginary work arrays
void rotate(int n, const cd *U, const cd *X, cd *Y, std::vector<double> &buf) {
    buf.resize((size_t)8 * n * n);
    double *ur = buf.data(), *ui = ur + n * n, *xr = ui + n * n, *xi = xr + n * n, *tr = xi + n * n,
           *ti = tr + n * n, *yr = ti + n * n, *yi = yr + n * n;
    for (int e = 0; e < n * n; ++e) {
        ur[e] = U[e].real();
        ui[e] = U[e].imag();
        xr[e] = X[e].real();
        xi[e] = X[e].imag();
        tr[e] = ti[e] = yr[e] = yi[e] = 0.0;
    }
    // T = U^+ X: T[i][j] = sum_p conj(U[p][i]) X[p][j]
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < n; ++p) {
            const double a = ur[p * n + i], b = ui[p * n + i];
            double *tri = tr + i * n, *tii = ti + i * n;
            const double *xrp = xr + p * n, *xip = xi + p * n;
            for (int j = 0; j < n; ++j) {
                tri[j] += a * xrp[j] + b * xip[j];
                tii[j] += a * xip[j] - b * xrp[j];
            }
        }
    // Y = T U
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < n; ++p) {
            const double a = tr[i * n + p], b = ti[i * n + p];
            double *yri = yr + i * n, *yii = yi + i * n;
            const double *urp = ur + p * n, *uip = ui + p * n;
            for (int j = 0; j < n; ++j) {
                yri[j] += a * urp[j] - b * uip[j];
                yii[j] += a * uip[j] + b * urp[j];
            }
        }
    for (int e = 0; e < n * n; ++e) Y[e] = cd(yr[e], yi[e]);
}

struct Params {
    int n, nw, gaussian, shift, inject;
    const double *omega;
    double width, mu, temperature, eta, window;
};

// One k point's contribution from its interpolated channels X[22][n][n]: acc_sc / acc_ic [nw][27].
void point_value(const Params &P, const cd *X, std::vector<cd> &work, std::vector<double> &buf, cd *acc_sc,
                 cd *acc_ic) {
    const int n = P.n, n2 = n * n;
    work.resize((size_t)n2 * 40);
    cd *Hf = work.data(), *U = Hf + n2, *R = U + n2;   // R: rotated channels 1..21 (21 n^2)
    cd *D = R + (size_t)21 * n2;                        // D[3]
    cd *Ab = D + (size_t)3 * n2;                        // Berry connection [3]
    std::vector<double> E;
    std::vector<cd> A(Hf, Hf + n2), V;
    // Hermitian matrix from the upper triangle (imaginary part of the diagonal ignored)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const cd h = j >= i ? X[(size_t)i * n + j] : std::conj(X[(size_t)j * n + i]);
            A[(size_t)i * n + j] = i == j ? cd(h.real(), 0.0) : h;
        }
    hermitian_eigen(n, A, E, V);
    std::copy(V.begin(), V.end(), U);
    // occupations and screening
    std::vector<double> f(n);
    const double kt = kBoltzmannEvK * P.temperature;
    for (int i = 0; i < n; ++i) {
        if (kt == 0.0) {
            f[i] = E[i] <= P.mu ? 1.0 : 0.0;
        } else {
            const double arg = std::min(700.0, std::max(-700.0, (E[i] - P.mu) / kt));
            f[i] = 1.0 / (std::exp(arg) + 1.0);
        }
    }
    const double w = P.width, cutoff = w * std::sqrt(200.0);
    std::vector<std::pair<int, int>> pairs;
    bool active_k = false;
    for (int nn = 0; nn < n; ++nn)
        for (int m = 0; m < n; ++m) {
            if (!(std::fabs(f[nn] - f[m]) > 1.0e-10)) continue;
            const double tr = E[m] - E[nn];
            bool in_pair = false, in_window = false;
            for (int q = 0; q < P.nw; ++q) {
                const double x = tr - P.omega[q];
                if (P.gaussian) {
                    if (std::fabs(x) < cutoff) in_pair = in_window = true;
                } else {
                    in_pair = true;
                    if (P.window <= 0.0 || std::fabs(x) < P.window * w) in_window = true;
                }
            }
            if (in_pair) pairs.emplace_back(nn, m);
            if (in_window) active_k = true;
        }
    if (!active_k || pairs.empty()) return;
    // rotated channels: 1..3 dH (V), 4..9 ddH (W, symmetric storage), 10..12 A, 13..21 dA
    for (int c = 1; c < kChannels; ++c) rotate(n, U, X + (size_t)c * n2, R + (size_t)(c - 1) * n2, buf);
    auto Vc = [&](int a) { return R + (size_t)a * n2; };
    auto Wc = [&](int a, int b) { return R + (size_t)(3 + sym_index(a, b)) * n2; };
    auto Ac = [&](int a) { return R + (size_t)(9 + a) * n2; };
    auto dAc = [&](int a, int b) { return R + (size_t)(12 + 3 * a + b) * n2; };
    std::vector<double> inv((size_t)n2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double de = E[i] - E[j];
            double v;
            if (P.eta == 0.0) {
                v = de == 0.0 ? 0.0 : 1.0 / de;
            } else {
                const double den = de * de + P.eta * P.eta;
                if (std::isfinite(den) && den != 0.0) {
                    v = de / den;
                } else {
                    const double s = std::max(std::fabs(de), P.eta), sd = de / s, sw = P.eta / s;
                    v = (sd / (sd * sd + sw * sw)) / s;
                }
            }
            inv[(size_t)i * n + j] = v;
        }
    for (int a = 0; a < 3; ++a)
        for (int e = 0; e < n2; ++e) {
            D[(size_t)a * n2 + e] = -Vc(a)[e] * inv[e];
            Ab[(size_t)a * n2 + e] = Ac(a)[e] + cd(0.0, 1.0) * D[(size_t)a * n2 + e];
        }
    auto at = [&](const cd *M, int i, int j) { return M[(size_t)i * n + j]; };
    // generalized derivative element (a = x, b = y of the reference's argument order)
    auto gderiv = [&](int x, int y, int nn, int m) {
        const cd *Ax = Ac(x), *Ay = Ac(y), *Vx = Vc(x), *Vy = Vc(y), *Dx = D + (size_t)x * n2, *Dy = D + (size_t)y * n2;
        cd sum_ad(0.0, 0.0), sum_hd(0.0, 0.0);
        for (int l = 0; l < n; ++l) {
            if (l == nn || l == m) continue;
            sum_ad += at(Ax, nn, l) * at(Dy, l, m) - at(Dy, nn, l) * at(Ax, l, m);
            sum_hd += at(Vx, nn, l) * at(Dy, l, m) - at(Dy, nn, l) * at(Vx, l, m);
        }
        const cd ad_bit = (at(Ay, nn, nn) - at(Ay, m, m)) * at(Dx, nn, m) + (at(Ax, nn, nn) - at(Ax, m, m)) * at(Dy, nn, m);
        const cd aa_bit = (at(Ay, nn, nn) - at(Ay, m, m)) * at(Ax, nn, m);
        const cd dv_bit = (at(Vy, nn, nn) - at(Vy, m, m)) * at(Dx, nn, m) + (at(Vx, nn, nn) - at(Vx, m, m)) * at(Dy, nn, m);
        return at(dAc(x, y), nn, m) + ad_bit - cd(0.0, 1.0) * aa_bit + sum_ad +
               cd(0.0, 1.0) * (at(Wc(x, y), nn, m) + sum_hd + dv_bit) * (-inv[(size_t)nn * n + m]);
    };
    const double gnorm = 1.0 / (std::sqrt(kPi) * w);
    std::vector<double> weight(P.nw);
    cd Gnm[3][3], Gmn[3][3], Qs[27], Qi[27];
    for (const auto &pr : pairs) {
        const int nn = pr.first, m = pr.second;
        const double df = f[nn] - f[m], tr = E[m] - E[nn];
        for (int q = 0; q < P.nw; ++q) {
            const double x = tr - P.omega[q];
            double dl;
            if (P.gaussian)
                dl = std::fabs(x) < cutoff ? gnorm * std::exp(-(x / w) * (x / w)) : 0.0;
            else
                dl = (w / kPi) / (x * x + w * w);
            weight[q] = df * dl;
        }
        if (P.shift) {
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) {
                    Gnm[x][y] = gderiv(x, y, nn, m);
                    Gmn[x][y] = gderiv(x, y, m, nn);
                }
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                        Qs[9 * a + 3 * b + c] = cd(0.0, -1.0) * (Gnm[c][a] * at(Ab + (size_t)b * n2, m, nn) -
                                                                 Gmn[b][a] * at(Ab + (size_t)c * n2, nn, m));
        }
        if (P.inject) {
            for (int a = 0; a < 3; ++a) {
                const cd vd = at(Vc(a), m, m) - at(Vc(a), nn, nn);
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                        Qi[9 * a + 3 * b + c] = vd * at(Ab + (size_t)c * n2, nn, m) * at(Ab + (size_t)b * n2, m, nn);
            }
        }
        for (int q = 0; q < P.nw; ++q) {
            if (weight[q] == 0.0) continue;
            for (int e = 0; e < 27; ++e) {
                if (P.shift) acc_sc[(size_t)q * 27 + e] += weight[q] * Qs[e];
                if (P.inject) acc_ic[(size_t)q * 27 + e] += weight[q] * Qi[e];
            }
        }
    }
}

}  // namespace

extern "C" {

// Sums over the mesh points of k3 planes [plane_begin, plane_end) of a Gamma-centred mesh (k_i = i / N_i,
// axis 1 fastest).  hr: [nr][n][n] complex (interleaved re, im); ar: [nr][3][n][n] or null; lattice rows
// a1, a2, a3.  Writes sc_out / ic_out [nw][27] complex (interleaved).  threads <= 0: OpenMP default.
// loop_seconds (nullable): wall time of the k-loop alone (after the R-space channels are assembled).
// Returns 0, or -1 when an R vector falls outside a box of span <= 64 per axis.
int oracle_kloop_planes(int n, int nr, const int *rvec, const int *deg, const double *lattice, const double *hr,
                        const double *ar, const int *mesh, int plane_begin, int plane_end, const double *omega, int nw,
                        double width, int gaussian, double mu, double temperature, double eta, double window,
                        int shift, int inject, int threads, double *sc_out, double *ic_out, double *loop_seconds) {
    const int n2 = n * n;
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int r = 0; r < nr; ++r)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], rvec[3 * r + a]);
            hi[a] = std::max(hi[a], rvec[3 * r + a]);
        }
    const int S[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    if (S[0] > 64 || S[1] > 64 || S[2] > 64) return -1;
    const size_t blk = (size_t)kChannels * n2;   // channels of one R vector
    // R box [S1][S2][S3][22][n][n] of channel rows with the degeneracy folded in (zero where no R)
    std::vector<cd> box((size_t)S[0] * S[1] * S[2] * blk, cd(0.0, 0.0));
    for (int r = 0; r < nr; ++r) {
        const double rc[3] = {rvec[3 * r] * lattice[0] + rvec[3 * r + 1] * lattice[3] + rvec[3 * r + 2] * lattice[6],
                              rvec[3 * r] * lattice[1] + rvec[3 * r + 1] * lattice[4] + rvec[3 * r + 2] * lattice[7],
                              rvec[3 * r] * lattice[2] + rvec[3 * r + 1] * lattice[5] + rvec[3 * r + 2] * lattice[8]};
        const size_t cell = (((size_t)(rvec[3 * r] - lo[0]) * S[1] + (rvec[3 * r + 1] - lo[1])) * S[2] + (rvec[3 * r + 2] - lo[2]));
        cd *out = box.data() + cell * blk;
        const double g = 1.0 / deg[r];
        const cd I(0.0, 1.0);
        for (int e = 0; e < n2; ++e) {
            const cd h(hr[2 * ((size_t)r * n2 + e)], hr[2 * ((size_t)r * n2 + e) + 1]);
            out[e] += g * h;
            for (int a = 0; a < 3; ++a) out[(size_t)(1 + a) * n2 + e] += g * I * rc[a] * h;
            for (int a = 0; a < 3; ++a)
                for (int b = a; b < 3; ++b) out[(size_t)(4 + sym_index(a, b)) * n2 + e] += -g * rc[a] * rc[b] * h;
            if (ar) {
                for (int a = 0; a < 3; ++a) {
                    const size_t s = ((size_t)r * 3 + a) * n2 + e;
                    const cd av(ar[2 * s], ar[2 * s + 1]);
                    out[(size_t)(10 + a) * n2 + e] += g * av;
                    for (int b = 0; b < 3; ++b) out[(size_t)(13 + 3 * a + b) * n2 + e] += g * I * rc[b] * av;
                }
            }
        }
    }
    const int N1 = mesh[0], N2 = mesh[1], N3 = mesh[2];
    Params P{n, nw, gaussian, shift, inject, omega, width, mu, temperature, eta, window};
    if (threads <= 0) threads = omp_get_max_threads();
    std::vector<std::vector<cd>> acc_sc(threads, std::vector<cd>((size_t)nw * 27)), acc_ic = acc_sc;
    std::vector<cd> y3((size_t)S[0] * S[1] * blk);
    const double t0 = omp_get_wtime();
    for (int k3 = plane_begin; k3 < plane_end; ++k3) {
        // R3 sum of the plane: y3[i1][i2] = sum_i3 e^{2 pi i R3 k3} box[i1][i2][i3]
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int ij = 0; ij < S[0] * S[1]; ++ij) {
            cd *o = y3.data() + (size_t)ij * blk;
            std::fill(o, o + blk, cd(0.0, 0.0));
            for (int i3 = 0; i3 < S[2]; ++i3) {
                const double ph = 2.0 * kPi * (double)(lo[2] + i3) * k3 / N3;
                const cd e3(std::cos(ph), std::sin(ph));
                const cd *x = box.data() + ((size_t)ij * S[2] + i3) * blk;
                for (size_t e = 0; e < blk; ++e) o[e] += e3 * x[e];
            }
        }
#pragma omp parallel num_threads(threads)
        {
            const int tid = omp_get_thread_num();
            std::vector<cd> y23((size_t)S[0] * blk), xk(blk), work;
            std::vector<double> buf;
#pragma omp for schedule(static)
            for (int k2 = 0; k2 < N2; ++k2) {
                for (int i1 = 0; i1 < S[0]; ++i1) {
                    cd *o = y23.data() + (size_t)i1 * blk;
                    std::fill(o, o + blk, cd(0.0, 0.0));
                    for (int i2 = 0; i2 < S[1]; ++i2) {
                        const double ph = 2.0 * kPi * (double)(lo[1] + i2) * k2 / N2;
                        const cd e2(std::cos(ph), std::sin(ph));
                        const cd *x = y3.data() + ((size_t)i1 * S[1] + i2) * blk;
                        for (size_t e = 0; e < blk; ++e) o[e] += e2 * x[e];
                    }
                }
                for (int k1 = 0; k1 < N1; ++k1) {
                    std::fill(xk.begin(), xk.end(), cd(0.0, 0.0));
                    for (int i1 = 0; i1 < S[0]; ++i1) {
                        const double ph = 2.0 * kPi * (double)(lo[0] + i1) * k1 / N1;
                        const cd e1(std::cos(ph), std::sin(ph));
                        const cd *x = y23.data() + (size_t)i1 * blk;
                        for (size_t e = 0; e < blk; ++e) xk[e] += e1 * x[e];
                    }
                    point_value(P, xk.data(), work, buf, acc_sc[tid].data(), acc_ic[tid].data());
                }
            }
        }
    }
    if (loop_seconds) *loop_seconds = omp_get_wtime() - t0;
    for (size_t e = 0; e < (size_t)nw * 27; ++e) {
        cd s(0.0, 0.0), t(0.0, 0.0);
        for (int th = 0; th < threads; ++th) {
            s += acc_sc[th][e];
            t += acc_ic[th][e];
        }
        if (sc_out) {
            sc_out[2 * e] = s.real();
            sc_out[2 * e + 1] = s.imag();
        }
        if (ic_out) {
            ic_out[2 * e] = t.real();
            ic_out[2 * e + 1] = t.imag();
        }
    }
    return 0;
}

// Eigensystem of the upper triangle of one Hermitian matrix (for tests): h [n][n] complex interleaved;
// e [n] ascending; v [n][n] columns (canonical phase).
int oracle_hermitian_eigen(int n, const double *h, double *e, double *v) {
    std::vector<cd> A((size_t)n * n), V;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const size_t s = j >= i ? (size_t)i * n + j : (size_t)j * n + i;
            cd x(h[2 * s], h[2 * s + 1]);
            if (j < i) x = std::conj(x);
            A[(size_t)i * n + j] = i == j ? cd(x.real(), 0.0) : x;
        }
    std::vector<double> E;
    hermitian_eigen(n, A, E, V);
    for (int i = 0; i < n; ++i) e[i] = E[i];
    for (size_t k = 0; k < (size_t)n * n; ++k) {
        v[2 * k] = V[k].real();
        v[2 * k + 1] = V[k].imag();
    }
    return 0;
}

}  // extern "C"
