// Berry-curvature dipole (BCD, Fermi-surface form) and intrinsic anomalous Hall (AHC, Matrix
// method) per-k stage, fused with the per-chunk k sum.
//
// Input per k point: the Hamiltonian-gauge channels written by the rotation kernel (dH_a, and with a
// position operator A_a and dA_ab = U^+ (d_b A_a) U) and the ascending eigenvalues.
//
// BCD (reference src/Responses/QuantumGeometry/ConventionalQuantumGeometry.jl:167-181,
//      BerryMultipoleMoments.jl:7-45, BerryMultipoleGeometry.jl:67-161):
//   Berry connection  B^a_nm = A^a_nm + i D^a_nm,  D^a_nm = -dH^a_nm inv(E_n - E_m)
//   Omega^n_ab = 2 sum_{m != n} -Im(B^a_nm B^b_mn) + Re F^H_nn,ab
//   Re F^H_nn,ab = Re(dA_ba - dA_ab)_nn + sum_m Im(A^a_nm A^b_mn - A^b_nm A^a_mn)
//   D[mu, a, b, c] += -f'(E_n; mu, T) Omega^n_ab Re dH^c_nn
//   every adjacent level pair must be separated by more than gap_tol
//   (OrthogonalBandDerivatives.jl:66-82).
// AHC (IntrinsicHallWorkspace.jl:6-23, IntrinsicHallGeometry.jl:117-136,
//      IntrinsicHallCurvature.jl:2-29, IntrinsicHallOccupations.jl:8-44):
//   occ = #{E_n <= mu};  v = dH + i (E_l - E_m) A;  r_lm = -i v_lm / (E_l - E_m) for l < occ <= m
//   and the reverse;  Omega_ab = sum_{l < occ} [Re F^H_ll,ab - sum_{m >= occ} Im(r^a_lm r^b_ml - r^b_lm r^a_ml)].
//
// One 256-thread workgroup per chunk of consecutive k points; the chunk sum runs over its k points in
// order inside the workgroup, so results do not depend on scheduling.
#include "wnlqg_common.h"
#include "wnlqg_kernels.h"

#include <math.h>

#define WN_GEOM_TPB 256
#define WN_GEOM_OUT 9  // outputs per thread: ceil((27 + 9) * WN_GEOM_MAXMU / 256)

namespace {

constexpr double kKbEv = 8.617333262145e-5;  // ResponseOccupations.jl:5

// -f'(E) in 1/eV (ResponseOccupations.jl:18-29): thermal at T > 0, Gaussian of width eta_fs at T = 0.
__device__ inline double wn_fs_weight(double e, double mu, double temperature, double eta_fs) {
    if (temperature > 0.0) {
        const double scale = kKbEv * temperature;
        const double z = exp(-fabs((e - mu) / scale));
        return z / (scale * (1.0 + z) * (1.0 + z));
    }
    const double x = (e - mu) / eta_fs;
    return exp(-x * x) / (sqrt(M_PI) * eta_fs);
}

__device__ inline double im_mul(double2 a, double2 b) { return a.x * b.y + a.y * b.x; }

// sum of v over the workgroup's `groups` row groups for the n * 9 (band, ab) entries; red holds
// [group][band][9]; result in red[band * 9 + ab] of group 0 (callers read it after the barrier).
__device__ inline void group_reduce9(double *red, const double v[9], int band, int grp, int groups, int n) {
    const int t = threadIdx.x;
    if (grp < groups) {
#pragma unroll
        for (int ab = 0; ab < 9; ++ab) red[(grp * n + band) * 9 + ab] = v[ab];
    }
    __syncthreads();
    for (int e = t; e < n * 9; e += WN_GEOM_TPB) {
        double s = red[e];
        for (int g = 1; g < groups; ++g) s += red[g * n * 9 + e];
        red[e] = s;
    }
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(WN_GEOM_TPB) void k_geom_f64(const double *__restrict__ x, const double *__restrict__ eg,
                                                         const double *__restrict__ mu, const int4 *__restrict__ chunks,
                                                         const double *__restrict__ wts, double2 *__restrict__ partial,
                                                         int *__restrict__ flags, long long kbase, WnShape sh,
                                                         WnGeom gp) {
    extern __shared__ double lds[];
    const int n = sh.n, n2 = n * n;
    const int t = threadIdx.x;
    const int nmu = gp.nmu;
    const bool do_bcd = (gp.kind & 1) != 0, do_ahc = (gp.kind & 2) != 0;
    // outputs: BCD [27][nmu] first, then AHC [9][nmu]
    const int aoff = do_bcd ? 27 * nmu : 0;
    const int nelem = aoff + (do_ahc ? 9 * nmu : 0);
    // LDS carve: E[64] | inv[n2] | Om[n*9] | Fd[n*9] | Vd[n*3] | W[nmu*n] | red[256*9] | res[9*nmu]
    double *E = lds;
    double *inv = E + 64;
    double *Om = inv + n2;
    double *Fd = Om + n * 9;
    double *Vd = Fd + n * 9;
    double *W = Vd + n * 3;
    double *red = W + (do_bcd ? nmu * n : 0);
    double *res = red + WN_GEOM_TPB * 9;

    const int groups = WN_GEOM_TPB / n;
    const int band = t % n, grp = t / n;
    const bool has_pos = gp.c_a >= 0;

    double acc[WN_GEOM_OUT];
#pragma unroll
    for (int j = 0; j < WN_GEOM_OUT; ++j) acc[j] = 0.0;

    const int4 ch = chunks[blockIdx.x];
    for (int kk = 0; kk < ch.y; ++kk) {
        const long long k = (long long)ch.x + kk;
        const double2 *X = (const double2 *)(x + (size_t)k * sh.ldo);
        const double2 *DH = X + (size_t)gp.c_dh * n2;
        const double2 *AA = has_pos ? X + (size_t)gp.c_a * n2 : nullptr;
        const double wk = wts ? wts[k] : 1.0;
        if (t < n) E[t] = eg[(size_t)k * n + t];
        __syncthreads();
        if (do_bcd && t + 1 < n && !(fabs(E[t + 1] - E[t]) > gp.gap_tol)) {
            atomicOr(flags + 2, 1);
            atomicMin(flags + 3, (int)(kbase + k));
        }
        for (int e = t; e < n2; e += WN_GEOM_TPB) {
            const int i = e / n, j = e - (e / n) * n;
            const double d = E[i] - E[j];
            inv[e] = gp.eta == 0.0 ? (d == 0.0 ? 0.0 : 1.0 / d) : wn_reg_inverse(d, gp.eta);
        }
        __syncthreads();

        // per band: connection products (BCD) and the commutator part of Re F^H_nn
        double om[9], fd[9];
#pragma unroll
        for (int ab = 0; ab < 9; ++ab) {
            om[ab] = 0.0;
            fd[ab] = 0.0;
        }
        if (grp < groups) {
            for (int m = grp; m < n; m += groups) {
                double2 an[3], am[3], bn[3], bm[3];
                const double ivn = inv[band * n + m], ivm = inv[m * n + band];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double2 dn = DH[(size_t)a * n2 + band * n + m];
                    const double2 dm = DH[(size_t)a * n2 + m * n + band];
                    an[a] = has_pos ? AA[(size_t)a * n2 + band * n + m] : make_double2(0.0, 0.0);
                    am[a] = has_pos ? AA[(size_t)a * n2 + m * n + band] : make_double2(0.0, 0.0);
                    // B = A + i D with D = -dH inv:  i D = (dH.y inv, -dH.x inv)
                    bn[a] = make_double2(an[a].x + dn.y * ivn, an[a].y - dn.x * ivn);
                    bm[a] = make_double2(am[a].x + dm.y * ivm, am[a].y - dm.x * ivm);
                }
                if (do_bcd && m != band) {
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) om[3 * a + b] -= im_mul(bn[a], bm[b]);
                }
                if (has_pos) {
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) fd[3 * a + b] += im_mul(an[a], am[b]) - im_mul(an[b], am[a]);
                }
            }
        }
        group_reduce9(red, fd, band, grp, groups, n);
        for (int e = t; e < n * 9; e += WN_GEOM_TPB) {
            const int bb = e / 9, ab = e - (e / 9) * 9, a = ab / 3, b = ab - (ab / 3) * 3;
            double f = red[e];
            if (has_pos) {
                const double2 *DA = X + (size_t)gp.c_da * n2;
                f += DA[(size_t)(3 * b + a) * n2 + bb * n + bb].x - DA[(size_t)(3 * a + b) * n2 + bb * n + bb].x;
            }
            Fd[e] = f;
        }
        __syncthreads();
        if (do_bcd) {
            group_reduce9(red, om, band, grp, groups, n);
            for (int e = t; e < n * 9; e += WN_GEOM_TPB) Om[e] = 2.0 * red[e] + Fd[e];
            if (t < n * 3) {
                const int bb = t / 3, c = t - (t / 3) * 3;
                Vd[t] = DH[(size_t)c * n2 + bb * n + bb].x;
            }
            for (int e = t; e < nmu * n; e += WN_GEOM_TPB) {
                const int q = e / n, bb = e - (e / n) * n;
                W[e] = wn_fs_weight(E[bb], mu[q], gp.temperature, gp.eta_fs);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < WN_GEOM_OUT; ++j) {
                const int o = t + j * WN_GEOM_TPB;
                if (o < aoff) {
                    // output component a + 3b + 9c (column-major [mu, a, b, c]); Om is stored [band][3a + b]
                    const int comp = o / nmu, q = o - (o / nmu) * nmu;
                    const int a = comp % 3, b = (comp / 3) % 3, c = comp / 9;
                    const int ab = 3 * a + b;
                    double s = 0.0;
                    for (int bb = 0; bb < n; ++bb) s += W[q * n + bb] * (Om[bb * 9 + ab] * Vd[bb * 3 + c]);
                    acc[j] += wk * s;
                }
            }
        }
        if (do_ahc) {
            for (int q = 0; q < nmu; ++q) {
                int occ = 0;
                for (int bb = 0; bb < n; ++bb) occ += E[bb] <= mu[q];
                const int nu = n - occ;
                double s[9];
#pragma unroll
                for (int ab = 0; ab < 9; ++ab) s[ab] = 0.0;
                for (int p = t; p < occ * nu; p += WN_GEOM_TPB) {
                    const int l = p / nu, m = occ + (p - (p / nu) * nu);
                    const double glm = E[l] - E[m], gml = E[m] - E[l];
                    double2 rl[3], rm[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const double2 dl = DH[(size_t)a * n2 + l * n + m];
                        const double2 dm = DH[(size_t)a * n2 + m * n + l];
                        const double2 al = has_pos ? AA[(size_t)a * n2 + l * n + m] : make_double2(0.0, 0.0);
                        const double2 amm = has_pos ? AA[(size_t)a * n2 + m * n + l] : make_double2(0.0, 0.0);
                        // v = dH + i gap A, then r = -i (v / gap)
                        const double2 vl = make_double2(dl.x - glm * al.y, dl.y + glm * al.x);
                        const double2 vm = make_double2(dm.x - gml * amm.y, dm.y + gml * amm.x);
                        const double2 ql = make_double2(vl.x / glm, vl.y / glm);
                        const double2 qm = make_double2(vm.x / gml, vm.y / gml);
                        rl[a] = make_double2(ql.y, -ql.x);
                        rm[a] = make_double2(qm.y, -qm.x);
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) s[3 * a + b] += im_mul(rl[a], rm[b]) - im_mul(rl[b], rm[a]);
                }
                // workgroup sum of the 9 link sums
#pragma unroll
                for (int ab = 0; ab < 9; ++ab) red[t * 9 + ab] = s[ab];
                __syncthreads();
                if (t < 9) {
                    double v = 0.0;
                    for (int l = 0; l < occ; ++l) v += Fd[l * 9 + t];
                    double links = 0.0;
                    for (int u = 0; u < WN_GEOM_TPB; ++u) links += red[u * 9 + t];
                    // t = 3a + b -> output component a + 3b (column-major [mu, a, b])
                    res[((t % 3) * 3 + t / 3) * nmu + q] = v - links;
                }
                __syncthreads();
            }
#pragma unroll
            for (int j = 0; j < WN_GEOM_OUT; ++j) {
                const int o = t + j * WN_GEOM_TPB;
                if (o >= aoff && o < nelem) acc[j] += wk * res[o - aoff];
            }
        }
        __syncthreads();
    }
    bool bad = false;
#pragma unroll
    for (int j = 0; j < WN_GEOM_OUT; ++j) {
        const int o = t + j * WN_GEOM_TPB;
        if (o < nelem) {
            bad |= !isfinite(acc[j]);
            partial[(size_t)blockIdx.x * nelem + o] = make_double2(acc[j], 0.0);
        }
    }
    if (bad) atomicOr(flags + 2, 2);
}

hipError_t wn_launch_geom_f64(const double *x, const double *e, const double *mu, const int4 *chunks, int nchunks,
                              const double *wts, double2 *partial, int *flags, long long kbase, const WnShape &sh,
                              const WnGeom &gp, hipStream_t st) {
    if (nchunks == 0) return hipSuccess;
    if (sh.n > 64 || gp.nmu < 1 || gp.nmu > WN_GEOM_MAXMU || gp.kind < 1 || gp.kind > 3)
        return hipErrorInvalidValue;
    const int n = sh.n;
    const size_t lds = sizeof(double) * (64 + (size_t)n * n + (size_t)n * 9 * 2 + (size_t)n * 3 +
                                         ((gp.kind & 1) ? (size_t)gp.nmu * n : 0) + (size_t)WN_GEOM_TPB * 9 +
                                         (size_t)9 * gp.nmu);
    hipLaunchKernelGGL(k_geom_f64, dim3(nchunks), dim3(WN_GEOM_TPB), lds, st, x, e, mu, chunks, wts, partial, flags,
                       kbase, sh, gp);
    return hipGetLastError();
}
