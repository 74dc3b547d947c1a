// HIP kernels of the MI355X k-mesh response engine (gfx950 / CDNA4, wave64).
//
// Stage map (reference functions in brackets):
//   build_coeff   R-space channels i R_a H, -R_a R_b H, A_a, i R_b A_a / deg(R) in MFMA fragment order
//                 [prepare_hamiltonian_real_space!, HamiltonianMatrixElements.jl:92-177;
//                  prepare_position_real_space!, PositionMatrixElements.jl:8-84]
//   phase_tables  exp(2 pi i k_a R_a) per axis   [_prepare_fourier_factors!, HamiltonianMatrixElements.jl:28-39]
//   interp_f64    X(k) = sum_R f_R(k) X(R) as a k x R GEMM on FP64 MFMA
//                 [_fourier_hamiltonian!, :42-58; derivative sums :427-462; PositionMatrixElements.jl:239-409]
//   eig_f64       Hermitian Jacobi, upper triangle, canonical column phase
//                 [DenseSpectra.hermitian_spectrum, DenseSpectra.jl:14-16; canonicalize_eigenvectors!, :6-25]
//   rotate_f64    U^+ X U                                [transform_to_hamiltonian_gauge!, :61-71]
//   pairs_f64     occupations, delta screen, active pairs, D, Berry connection, generalized derivative,
//                 SC/IC kernels   [TransitionScreening.jl:32-162; PositionMatrixElements.jl:444-489;
//                 ConventionalShiftCurrent.jl:191-229,527-767; ConventionalInjectionCurrent.jl:23-97]
//   accum         sum over pairs of df * delta(w) * Q per chunk of k points
//                 [accumulate_conventional_response!, ConventionalShiftCurrent.jl:774-884]
//   lane_reduce   chunk partials -> deterministic lanes [consume_point!, IntegralDriver.jl:820-874]

#include <type_traits>

#include "wnlqg_common.h"
#include "wnlqg_kernels.h"

#include <math.h>

#include <algorithm>
#include <stdlib.h>

// ------------------------------------------------------------------------------------------------
// R-space channels in MFMA fragment order.
//
// Real GEMM: out[k][2J+p] = sum_kk A[k][kk] B[kk][2J+p] with, per block of 8 R vectors, 16 K rows
// ordered (cos part of r0..r7, sin part of r0..r7).  B entries for complex X_J(R)/deg(R):
//   cos row: (Re X, Im X) into (re, im) output columns; sin row: (-Im X, Re X).
// Fragment order: bf[((kb*nNt + nt)*4 + s)*8 + j][lane] holds B[kb*16 + 4s + lane/16][nt*128 + 16j + lane%16],
// which is the per-lane B operand of v_mfma_f64_16x16x4f64 for k-substep s and 16-column subtile j.
// ------------------------------------------------------------------------------------------------
// B entry of R vector r, K part (0: multiplies cos, 1: multiplies sin) and real output column col.
__device__ double wn_coeff_entry(const double *__restrict__ hr, const double *__restrict__ ar,
                                 const double *__restrict__ rc, const int *__restrict__ deg, const WnShape &sh, int r,
                                 int part, int col) {
    const int n = sh.n, n2 = n * n;
    const int J = col >> 1, p2 = col & 1;
    if (r >= sh.nr || J >= sh.C * n2) return 0.0;
    const int slot = J / n2, ij = J - slot * n2;
    const int c = sh.lch[slot];
    const int i = ij / n, jj = ij - i * n;
    // Julia column-major element (i, jj, r) and (i, jj, a, r)
    const double *h = hr + 2 * ((size_t)i + (size_t)n * jj + (size_t)n2 * r);
    double xr = 0.0, xi = 0.0;
    const double *R = rc + 3 * r;
    if (c == WN_CH_H) {
        xr = h[0];
        xi = h[1];
    } else if (c < WN_CH_DDH) {
        const double d = R[c - WN_CH_DH];
        // (i d) * h
        xr = -d * h[1];
        xi = d * h[0];
    } else if (c < WN_CH_A) {
        const int t = c - WN_CH_DDH;
        const int a = t < 3 ? t : (t == 3 ? 1 : 0);
        const int b = t < 3 ? t : (t == 5 ? 1 : 2);
        const double f = -R[a] * R[b];
        xr = f * h[0];
        xi = f * h[1];
    } else {
        const int a = c < WN_CH_DA ? c - WN_CH_A : (c - WN_CH_DA) / 3;
        const double *p = ar + 2 * ((size_t)i + (size_t)n * jj + (size_t)n2 * a + (size_t)3 * n2 * r);
        if (c < WN_CH_DA) {
            xr = p[0];
            xi = p[1];
        } else {
            const int b = (c - WN_CH_DA) % 3;
            const double d = R[b];
            xr = -d * p[1];
            xi = d * p[0];
        }
    }
    const double inv_deg = 1.0 / (double)deg[r];
    xr *= inv_deg;
    xi *= inv_deg;
    return part == 0 ? (p2 == 0 ? xr : xi) : (p2 == 0 ? -xi : xr);
}

__global__ void k_build_coeff(const double *__restrict__ hr, const double *__restrict__ ar,
                              const double *__restrict__ rc, const int *__restrict__ deg, double *__restrict__ bf,
                              WnShape sh) {
    const long long total = (long long)sh.Kb * sh.nNt * 2048;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int lane = (int)(e & 63);
        const int j = (int)((e >> 6) & 7);
        const int s = (int)((e >> 9) & 3);
        const long long rest = e >> 11;
        const int nt = (int)(rest % sh.nNt);
        const int kb = (int)(rest / sh.nNt);
        const int kk = 4 * s + (lane >> 4);
        const int part = kk >> 3;
        const int r = kb * WN_RB + (kk & 7);
        const int col = nt * WN_GN + 16 * j + (lane & 15);
        bf[e] = wn_coeff_entry(hr, ar, rc, deg, sh, r, part, col);
    }
}

// Split-BF16 operand of v_mfma_f32_32x32x16_bf16 in fragment order: element
// e = ((kb*nT32 + ct)*64 + lane)*8 + j holds B[kk = 8*(lane/32) + j][32*ct + lane%32] of K block kb, where
// within a block kk = 2*rr + part (R vector kb*8 + rr, cos/sin part).  x = hi + mid + lo with each part the
// BF16 rounding of the remainder.
__global__ void k_build_coeff_bf16(const double *__restrict__ hr, const double *__restrict__ ar,
                                   const double *__restrict__ rc, const int *__restrict__ deg,
                                   __bf16 *__restrict__ bh, __bf16 *__restrict__ bm, __bf16 *__restrict__ bl,
                                   WnShape sh) {
    const int nT32 = sh.ldo / 32;
    const long long total = (long long)sh.Kb * nT32 * 512;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e & 7);
        const int lane = (int)((e >> 3) & 63);
        const long long rest = e >> 9;
        const int ct = (int)(rest % nT32);
        const int kb = (int)(rest / nT32);
        const int kk = 8 * (lane >> 5) + j;
        const int r = kb * WN_RB + (kk >> 1);
        const int part = kk & 1;
        const int col = 32 * ct + (lane & 31);
        const double x = wn_coeff_entry(hr, ar, rc, deg, sh, r, part, col);
        const float xf = (float)x;
        const __bf16 h = (__bf16)xf;
        const double r1 = x - (double)(float)h;
        const __bf16 m = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)m;
        bh[e] = h;
        bm[e] = m;
        if (bl) bl[e] = (__bf16)(float)r2;
    }
}

hipError_t wn_launch_build_coeff_bf16(const double *hr, const double *ar, const double *rc, const int *deg,
                                      __bf16 *bh, __bf16 *bm, __bf16 *bl, const WnShape &sh, hipStream_t st) {
    long long total = (long long)sh.Kb * (sh.ldo / 32) * 512;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_build_coeff_bf16, dim3((unsigned)blocks), dim3(256), 0, st, hr, ar, rc, deg, bh, bm, bl, sh);
    return hipGetLastError();
}

hipError_t wn_launch_build_coeff(const double *hr, const double *ar, const double *rc, const int *deg, double *bf,
                                 const WnShape &sh, hipStream_t st) {
    long long total = (long long)sh.Kb * sh.nNt * 2048;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_build_coeff, dim3((unsigned)blocks), dim3(256), 0, st, hr, ar, rc, deg, bf, sh);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Mesh k points and phase tables.
// ------------------------------------------------------------------------------------------------
__global__ void k_mesh_kpoints(double *__restrict__ kp, int n1, int n2, int n3, long long k0, int count, int rows) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows) return;
    double k1 = 0.0, k2 = 0.0, k3 = 0.0;
    if (t < count) {
        long long g = k0 + t;
        long long i1 = g % n1;
        long long i2 = (g / n1) % n2;
        long long i3 = g / ((long long)n1 * n2);
        k1 = (double)i1 / (double)n1;
        k2 = (double)i2 / (double)n2;
        k3 = (double)i3 / (double)n3;
    }
    kp[3 * t + 0] = k1;
    kp[3 * t + 1] = k2;
    kp[3 * t + 2] = k3;
}

hipError_t wn_launch_mesh_kpoints(double *kp, const int *mesh, long long k0, int count, int rows_pad,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_kpoints, dim3((rows_pad + 255) / 256), dim3(256), 0, st, kp, mesh[0], mesh[1], mesh[2],
                       k0, count, rows_pad);
    return hipGetLastError();
}

__global__ void k_phase_tables(const double *__restrict__ kp, int rmin0, int rmin1, int rmin2, double2 *__restrict__ tab,
                               float2 *__restrict__ tab32, int rows, int S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = rows * 3 * S;
    if (t >= total) return;
    int k = t / (3 * S);
    int rem = t - k * 3 * S;
    int a = rem / S;
    int j = rem - a * S;
    int rmin = a == 0 ? rmin0 : (a == 1 ? rmin1 : rmin2);
    double x = 2.0 * kp[3 * k + a] * (double)(rmin + j);
    double s, c;
    sincospi(x, &s, &c);
    tab[t] = make_double2(c, s);
    if (tab32) tab32[t] = make_float2((float)c, (float)s);
}

hipError_t wn_launch_phase_tables(const double *kp, const int *rmin, double2 *tab, float2 *tab32, int rows, int S,
                                  hipStream_t st) {
    int total = rows * 3 * S;
    hipLaunchKernelGGL(k_phase_tables, dim3((total + 255) / 256), dim3(256), 0, st, kp, rmin[0], rmin[1], rmin[2], tab,
                       tab32, rows, S);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// FP64 MFMA interpolation GEMM.
//
// Workgroup: 256 threads = 4 waves in 2x2, tile 128 k x 128 real columns; each wave owns a 64x64
// block = 4x4 subtiles of v_mfma_f64_16x16x4f64.  The A operand (Fourier factors) is generated in
// registers from the per-axis phase tables: lane l needs rows 16i + l%16 and R vectors
// kb*8 + l/16 and kb*8 + 4 + l/16 of each K block, cos parts for substeps 0,1 and sin parts for
// substeps 2,3.  B fragments stream from HBM/L2 in fragment order, one K block ahead.
// XCD-aware tile order: workgroup g runs on XCD g%8; each XCD sweeps all k tiles of its own
// column panels so that a 2.7 MB B panel stays in that XCD's L2.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_interp_f64(const double *__restrict__ bf, const double2 *__restrict__ tab,
                                                      const int4 *__restrict__ ridx, double *__restrict__ out,
                                                      int nMt, int nNt, int Kb, int S, int ldo) {
    const int g = blockIdx.x;
    const int xcd = g & 7;
    const int jj = g >> 3;
    const int nl = jj / nMt;
    const int mt = jj - nl * nMt;
    const int nt = nl * 8 + xcd;
    if (nt >= nNt) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 15, lq = lane >> 4;

    const double2 *tk[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) tk[i] = tab + (size_t)(mt * WN_GM + wm * 64 + 16 * i + lrow) * 3 * S;

    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};

    const size_t kb_stride = (size_t)nNt * 2048;
    const double *bp = bf + (size_t)nt * 2048 + (size_t)wn * 256 + lane;
    double bcur[4][4], bnext[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) bcur[s][j] = bp[s * 512 + j * 64];

    for (int kb = 0; kb < Kb; ++kb) {
        if (kb + 1 < Kb) {
            const double *bq = bp + (size_t)(kb + 1) * kb_stride;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int j = 0; j < 4; ++j) bnext[s][j] = bq[s * 512 + j * 64];
        }
        const int4 id0 = ridx[kb * WN_RB + lq];
        const int4 id1 = ridx[kb * WN_RB + 4 + lq];
        double a[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double2 *t = tk[i];
            double2 f0 = cmul(cmul(t[id0.x], t[S + id0.y]), t[2 * S + id0.z]);
            double2 f1 = cmul(cmul(t[id1.x], t[S + id1.y]), t[2 * S + id1.z]);
            a[0][i] = f0.x;
            a[1][i] = f1.x;
            a[2][i] = f0.y;
            a[3][i] = f1.y;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][i], bcur[s][j], acc[i][j], 0, 0, 0);
        if (kb + 1 < Kb) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int j = 0; j < 4; ++j) bcur[s][j] = bnext[s][j];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t row = (size_t)mt * WN_GM + wm * 64 + 16 * i + lq + 4 * q;
            double *o = out + row * ldo + (size_t)nt * WN_GN + wn * 64 + lrow;
#pragma unroll
            for (int j = 0; j < 4; ++j) o[16 * j] = acc[i][j][q];
        }
    }
}

hipError_t wn_launch_interp_f64(const double *bf, const double2 *tab, const int4 *ridx, double *out, int rows_pad,
                                const WnShape &sh, hipStream_t st) {
    const int nMt = rows_pad / WN_GM;
    const int panels = (sh.nNt + 7) / 8;
    const long long grid = 8LL * panels * nMt;
    hipLaunchKernelGGL(k_interp_f64, dim3((unsigned)grid), dim3(256), 0, st, bf, tab, ridx, out, nMt, sh.nNt, sh.Kb,
                       sh.S, sh.ldo);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Split-BF16 MFMA interpolation GEMM (mixed precision forms).
//
// Operands are split into BF16 parts, x ~ x_h + x_m (+ x_l), each part the BF16 rounding of the
// remainder, and products are accumulated by v_mfma_f32_32x32x16_bf16 from the terms
//   NT=3: hh + hm + mh   NT=4: + h*B_l   NT=5: + mm   NT=6: + A_l*h
// (A = Fourier factors, B = coefficients).  FP32 accumulators are flushed into FP64 accumulators
// every WN_FLUSH K blocks, so the FP32 rounding of a long K sum does not limit the result.
// Workgroup 256 threads = 2x2 waves, tile 128 k x 128 real columns, wave tile 64x64 = 2x2 subtiles
// of 32x32.  Per K block (8 R vectors) the workgroup generates the A fragments of its 128 rows once
// (thread t: row subtile t/64, lane t%64 -> row 32*(t/64) + lane%32, R vectors 4*(lane/32)+{0..3}
// as (cos, sin) pairs, in FP32 from the phase tables) into double-buffered LDS; B fragments stream
// from HBM/L2 in fragment order one K block ahead.  XCD-aware tile order as in the FP64 kernel.
// ------------------------------------------------------------------------------------------------
#define WN_FLUSH 4

__device__ __forceinline__ void wn_split_phase(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    l = (__bf16)(r1 - (float)m);
}

template <int NT>
__global__ __launch_bounds__(256) void k_interp_bf16(const bf16x8 *__restrict__ bh, const bf16x8 *__restrict__ bm,
                                                    const bf16x8 *__restrict__ bl, const float2 *__restrict__ tab,
                                                    const int4 *__restrict__ ridx, double *__restrict__ out, int nMt,
                                                    int nt0, int nNt1, int Kb, int S, int ldo) {
    constexpr int PA = NT >= 6 ? 3 : 2;  // parts of A in use
    constexpr int PB = NT >= 4 ? 3 : 2;  // parts of B in use
    __shared__ bf16x8 afrag[2][PA][4][64];
    const int g = blockIdx.x;
    const int xcd = g & 7;
    const int jj = g >> 3;
    const int nl = jj / nMt;
    const int mt = jj - nl * nMt;
    const int nt = nt0 + nl * 8 + xcd;
    if (nt >= nNt1) return;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l32 = lane & 31, lh = lane >> 5;
    const int nT32 = ldo / 32;

    // phase generation role of this thread
    const int grs = t >> 6;
    const float2 *tk = tab + (size_t)(mt * WN_GM + 32 * grs + l32) * 3 * S;
    auto gen = [&](int kb, int buf) {
        bf16x8 ph[PA];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int4 id = ridx[kb * WN_RB + 4 * lh + rr];
            const float2 t1 = tk[id.x], t2 = tk[S + id.y], t3 = tk[2 * S + id.z];
            const float pr = t1.x * t2.x - t1.y * t2.y, pi = t1.x * t2.y + t1.y * t2.x;
            const float fr = pr * t3.x - pi * t3.y, fi = pr * t3.y + pi * t3.x;
            __bf16 h0, m0, l0, h1, m1, l1;
            wn_split_phase(fr, h0, m0, l0);
            wn_split_phase(fi, h1, m1, l1);
            ph[0][2 * rr] = h0;
            ph[0][2 * rr + 1] = h1;
            ph[1][2 * rr] = m0;
            ph[1][2 * rr + 1] = m1;
            if (PA == 3) {
                ph[PA - 1][2 * rr] = l0;
                ph[PA - 1][2 * rr + 1] = l1;
            }
        }
#pragma unroll
        for (int p = 0; p < PA; ++p) afrag[buf][p][grs][lane] = ph[p];
    };

    f32x16 acc[2][2];
    double acc64[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                acc[i][j][q] = 0.0f;
                acc64[i][j][q] = 0.0;
            }

    const int ct0 = nt * 4 + wn * 2;
    bf16x8 bc[PB][2], bn[PB][2];
    auto loadb = [&](int kb, bf16x8 (&dst)[PB][2]) {
        const size_t boff = ((size_t)kb * nT32 + ct0) * 64 + lane;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            dst[0][j] = bh[boff + 64 * j];
            dst[1][j] = bm[boff + 64 * j];
            if (PB == 3) dst[PB - 1][j] = bl[boff + 64 * j];
        }
    };
    loadb(0, bc);
    gen(0, 0);
    __syncthreads();
    for (int kb = 0; kb < Kb; ++kb) {
        const int buf = kb & 1;
        if (kb + 1 < Kb) {
            loadb(kb + 1, bn);
            gen(kb + 1, buf ^ 1);
        }
        bf16x8 a[PA][2];
#pragma unroll
        for (int p = 0; p < PA; ++p)
#pragma unroll
            for (int i = 0; i < 2; ++i) a[p][i] = afrag[buf][p][2 * wm + i][lane];
        // small terms first
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (NT >= 6) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[PA - 1][i], bc[0][j], acc[i][j], 0, 0, 0);
                if (NT >= 5) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], bc[1][j], acc[i][j], 0, 0, 0);
                if (NT >= 4) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], bc[PB - 1][j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], bc[0][j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], bc[1][j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], bc[0][j], acc[i][j], 0, 0, 0);
        if ((kb % WN_FLUSH) == WN_FLUSH - 1 || kb + 1 == Kb) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        acc64[i][j][q] += (double)acc[i][j][q];
                        acc[i][j][q] = 0.0f;
                    }
        }
        if (kb + 1 < Kb) {
#pragma unroll
            for (int p = 0; p < PB; ++p)
#pragma unroll
                for (int j = 0; j < 2; ++j) bc[p][j] = bn[p][j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const size_t row = (size_t)mt * WN_GM + wm * 64 + 32 * i + (q >> 2) * 8 + lh * 4 + (q & 3);
            double *o = out + row * ldo + (size_t)nt * 128 + wn * 64 + l32;
#pragma unroll
            for (int j = 0; j < 2; ++j) o[32 * j] = acc64[i][j][q];
        }
}

hipError_t wn_launch_interp_bf16(const __bf16 *bh, const __bf16 *bm, const __bf16 *bl, const float2 *tab,
                                 const int4 *ridx, double *out, int rows_pad, int nterms, int col0, const WnShape &sh,
                                 hipStream_t st) {
    const int nMt = rows_pad / WN_GM;
    const int nNt1 = sh.ldo / 128;
    const int nt0 = col0 / 128;
    const int panels = (nNt1 - nt0 + 7) / 8;
    if (panels <= 0) return hipSuccess;
    const long long grid = 8LL * panels * nMt;
#define WN_BF16_CASE(T)                                                                                        \
    if (nterms == T) {                                                                                         \
        hipLaunchKernelGGL(k_interp_bf16<T>, dim3((unsigned)grid), dim3(256), 0, st, (const bf16x8 *)bh,        \
                           (const bf16x8 *)bm, (const bf16x8 *)bl, tab, ridx, out, nMt, nt0, nNt1, sh.Kb, sh.S,  \
                           sh.ldo);                                                                            \
        return hipGetLastError();                                                                              \
    }
    WN_BF16_CASE(3)
#undef WN_BF16_CASE
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------------------
// Split-FP16 MFMA interpolation GEMM (the mixed precision form).
//
// Both operands are split into two FP16 parts, x ~ x_h + x_m with x_h = fp16(x), x_m = fp16(x - x_h),
// i.e. ~22 significant bits.  Coefficient columns are pre-scaled by an exact power of two so that
// their largest entry lies in [2^14, 2^15) (FP16 normal range for everything that matters); the
// scale is removed when the result is written.  v_mfma_f32_32x32x16_f16 accumulates
//   NT=3: hh + hm + mh       NT=4: + mm
// in FP32 within a K block; FP32 accumulators are flushed into FP64 every WN_FLUSH K blocks.
// Workgroup 256 threads = 4 waves stacked in k, tile 128 k x 64 real columns, wave tile 32 x 64
// (1x2 subtiles): each wave generates the Fourier factors of its own 32 rows (no sharing, no LDS,
// no barriers) from the per-k (R1,R2)-plane table T12 and the R3 table T3, and all four waves read
// the same B fragments (L1 reuse).  XCD-aware tile order as in the FP64 kernel.
// ------------------------------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256, 2) void k_interp_f16(const f16x8 *__restrict__ bhp, const f16x8 *__restrict__ bmp,
                                                      const double *__restrict__ cscale,
                                                      const float2 *__restrict__ t12, const float2 *__restrict__ t3,
                                                      const int2 *__restrict__ ridx2, double *__restrict__ out,
                                                      int nMt, int nt0, int nNt64, int Kb, int S, int ldo) {
    const int g = blockIdx.x;
    const int xcd = g & 7;
    const int jj = g >> 3;
    const int nl = jj / nMt;
    const int mt = jj - nl * nMt;
    const int nt = nt0 + nl * 8 + xcd;
    if (nt >= nNt64) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int nT32 = ldo / 32;
    const int row = mt * WN_GM + wave * 32 + l32;
    const float2 *p12 = t12 + (size_t)row * S * S;
    const float2 *p3 = t3 + (size_t)row * S;

    f32x16 acc[2];
    double acc64[2][16];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            acc[j][q] = 0.0f;
            acc64[j][q] = 0.0;
        }
    const int ct0 = nt * 2;
    f16x8 bh[2], bm[2], nh[2], nm[2];
    {
        const size_t boff = (size_t)ct0 * 64 + lane;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            bh[j] = bhp[boff + 64 * j];
            bm[j] = bmp[boff + 64 * j];
        }
    }
    for (int kb = 0; kb < Kb; ++kb) {
        if (kb + 1 < Kb) {
            const size_t boff = ((size_t)(kb + 1) * nT32 + ct0) * 64 + lane;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                nh[j] = bhp[boff + 64 * j];
                nm[j] = bmp[boff + 64 * j];
            }
        }
        f16x8 ah, am;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int2 id = ridx2[kb * WN_RB + 4 * lh + rr];
            const float2 a = p12[id.x], b = p3[id.y];
            const float fr = a.x * b.x - a.y * b.y, fi = a.x * b.y + a.y * b.x;
            const _Float16 hr = (_Float16)fr, hi = (_Float16)fi;
            ah[2 * rr] = hr;
            ah[2 * rr + 1] = hi;
            am[2 * rr] = (_Float16)(fr - (float)hr);
            am[2 * rr + 1] = (_Float16)(fi - (float)hi);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (NT >= 4) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(am, bm[j], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(am, bh[j], acc[j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bm[j], acc[j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[j], acc[j], 0, 0, 0);
        if ((kb % WN_FLUSH) == WN_FLUSH - 1 || kb + 1 == Kb) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    acc64[j][q] += (double)acc[j][q];
                    acc[j][q] = 0.0f;
                }
        }
        if (kb + 1 < Kb) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                bh[j] = nh[j];
                bm[j] = nm[j];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = nt * 64 + 32 * j + l32;
        const double sc = cscale[col];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const size_t r = (size_t)mt * WN_GM + wave * 32 + (q >> 2) * 8 + lh * 4 + (q & 3);
            out[r * ldo + col] = acc64[j][q] * sc;
        }
    }
}

hipError_t wn_launch_interp_f16(const _Float16 *bh, const _Float16 *bm, const double *cscale, const float2 *t12,
                                const float2 *t3, const int2 *ridx2, double *out, int rows_pad, int nterms, int col0,
                                const WnShape &sh, hipStream_t st) {
    const int nMt = rows_pad / WN_GM;
    const int nNt64 = sh.ldo / 64;
    const int nt0 = col0 / 64;
    const int panels = (nNt64 - nt0 + 7) / 8;
    if (panels <= 0) return hipSuccess;
    const long long grid = 8LL * panels * nMt;
    if (nterms >= 4)
        hipLaunchKernelGGL(k_interp_f16<4>, dim3((unsigned)grid), dim3(256), 0, st, (const f16x8 *)bh,
                           (const f16x8 *)bm, cscale, t12, t3, ridx2, out, nMt, nt0, nNt64, sh.Kb, sh.S, sh.ldo);
    else
        hipLaunchKernelGGL(k_interp_f16<3>, dim3((unsigned)grid), dim3(256), 0, st, (const f16x8 *)bh,
                           (const f16x8 *)bm, cscale, t12, t3, ridx2, out, nMt, nt0, nNt64, sh.Kb, sh.S, sh.ldo);
    return hipGetLastError();
}

// Per-column power-of-two scale of the coefficient operand: largest |B| of column col into [2^14, 2^15).
__global__ void k_coeff_colscale(const double *__restrict__ hr, const double *__restrict__ ar,
                                 const double *__restrict__ rc, const int *__restrict__ deg, double *__restrict__ up,
                                 double *__restrict__ down, WnShape sh) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= sh.ldo) return;
    double m = 0.0;
    for (int r = 0; r < sh.nr; ++r)
        for (int part = 0; part < 2; ++part) m = fmax(m, fabs(wn_coeff_entry(hr, ar, rc, deg, sh, r, part, col)));
    int e = 0;
    if (m > 0.0) {
        frexp(m, &e);      // m = f * 2^e, f in [0.5, 1)
        e = 15 - e;        // m * 2^(15-e) in [2^14, 2^15)
    }
    up[col] = ldexp(1.0, e);
    down[col] = ldexp(1.0, -e);
}

// Split-FP16 operand of v_mfma_f32_32x32x16_f16 in the fragment order of k_build_coeff_bf16.
__global__ void k_build_coeff_f16(const double *__restrict__ hr, const double *__restrict__ ar,
                                  const double *__restrict__ rc, const int *__restrict__ deg,
                                  const double *__restrict__ up, _Float16 *__restrict__ bh, _Float16 *__restrict__ bm,
                                  WnShape sh) {
    const int nT32 = sh.ldo / 32;
    const long long total = (long long)sh.Kb * nT32 * 512;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e & 7);
        const int lane = (int)((e >> 3) & 63);
        const long long rest = e >> 9;
        const int ct = (int)(rest % nT32);
        const int kb = (int)(rest / nT32);
        const int kk = 8 * (lane >> 5) + j;
        const int r = kb * WN_RB + (kk >> 1);
        const int part = kk & 1;
        const int col = 32 * ct + (lane & 31);
        const double x = wn_coeff_entry(hr, ar, rc, deg, sh, r, part, col) * up[col];
        const _Float16 h = (_Float16)(float)x;
        bh[e] = h;
        bm[e] = (_Float16)(float)(x - (double)(float)h);
    }
}

hipError_t wn_launch_build_coeff_f16(const double *hr, const double *ar, const double *rc, const int *deg,
                                     double *up, double *down, _Float16 *bh, _Float16 *bm, const WnShape &sh,
                                     hipStream_t st) {
    hipLaunchKernelGGL(k_coeff_colscale, dim3((sh.ldo + 255) / 256), dim3(256), 0, st, hr, ar, rc, deg, up, down, sh);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    long long total = (long long)sh.Kb * (sh.ldo / 32) * 512;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_build_coeff_f16, dim3((unsigned)blocks), dim3(256), 0, st, hr, ar, rc, deg, up, bh, bm, sh);
    return hipGetLastError();
}

// Per-k tables for the FP16 path: T12[k][j1*S + j2] = T[k][0][j1] T[k][1][j2] (FP64 product rounded to FP32)
// and T3[k][j3] = T[k][2][j3].
__global__ void k_phase_planes(const double2 *__restrict__ tab, float2 *__restrict__ t12, float2 *__restrict__ t3,
                               int rows, int S) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = S * S + S;
    if (t >= rows * per) return;
    const int k = t / per, e = t - k * per;
    const double2 *tk = tab + (size_t)k * 3 * S;
    if (e < S * S) {
        const int j1 = e / S, j2 = e - j1 * S;
        const double2 v = cmul(tk[j1], tk[S + j2]);
        t12[(size_t)k * S * S + e] = make_float2((float)v.x, (float)v.y);
    } else {
        const double2 v = tk[2 * S + (e - S * S)];
        t3[(size_t)k * S + (e - S * S)] = make_float2((float)v.x, (float)v.y);
    }
}

hipError_t wn_launch_phase_planes(const double2 *tab, float2 *t12, float2 *t3, int rows, int S, hipStream_t st) {
    const int total = rows * (S * S + S);
    hipLaunchKernelGGL(k_phase_planes, dim3((total + 255) / 256), dim3(256), 0, st, tab, t12, t3, rows, S);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Separable interpolation on a uniform mesh.
//
// On the Gamma-centred mesh the Fourier factor factorizes, exp(2 pi i R.k) = e1(k1,R1) e2(k2,R2) e3(k3,R3),
// so X(k) = sum_R1 e1 sum_R2 e2 sum_R3 e3 X(R) is evaluated as three small contractions over a dense
// box of R vectors (absent R are zero), all in FP64:
//   A (per k3 plane):  Y3[R1R2][m]       = sum_R3 e3(k3,R3) X[R1R2R3][m]
//   B (per k2 range):  Y23[R1][k2][m]    = sum_R2 e2(k2,R2) Y3[R1R2][m]
//   C (per k point):   X[k2][k1][m]      = sum_R1 e1(k1,R1) Y23[R1][k2][m]
// m runs over the C*n^2 complex channel entries of a k row (the same row layout as the direct GEMM
// output).  Stage C writes every interpolated matrix once and is bound by HBM writes.
// ------------------------------------------------------------------------------------------------
__global__ void k_build_box(const double *__restrict__ hr, const double *__restrict__ ar, const double *__restrict__ rc,
                            const int *__restrict__ deg, const int *__restrict__ boxmap, double *__restrict__ box,
                            long long nbox, WnShape sh) {
    const long long total = nbox * sh.ldo;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long rb = e / sh.ldo;
        const int col = (int)(e - rb * sh.ldo);
        const int r = boxmap[rb];
        box[e] = r < 0 ? 0.0 : wn_coeff_entry(hr, ar, rc, deg, sh, r, 0, col);
    }
}

hipError_t wn_launch_build_box(const double *hr, const double *ar, const double *rc, const int *deg, const int *boxmap,
                               double *box, long long nbox, const WnShape &sh, hipStream_t st) {
    long long total = nbox * sh.ldo;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_build_box, dim3((unsigned)blocks), dim3(256), 0, st, hr, ar, rc, deg, boxmap, box, nbox, sh);
    return hipGetLastError();
}

// E[i][j] = exp(2 pi i (i/N) (rmin + j)) for one mesh axis (i/N as in integral_kpoint!).
__global__ void k_axis_phases(double2 *__restrict__ E, int N, int span, int rmin) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * span) return;
    const int i = t / span, j = t - i * span;
    const double k = (double)i / (double)N;
    double s, c;
    sincospi(2.0 * k * (double)(rmin + j), &s, &c);
    E[t] = make_double2(c, s);
}

hipError_t wn_launch_axis_phases(double2 *E, int N, int span, int rmin, hipStream_t st) {
    hipLaunchKernelGGL(k_axis_phases, dim3((N * span + 255) / 256), dim3(256), 0, st, E, N, span, rmin);
    return hipGetLastError();
}

// out[b*ob + p*op + m] = sum_{q<Q} E[(p0 + p)*Q + q] * in[b*ib + q*iq + m]   (complex, m < M, p < P)
// n2c32 > 0: entries m >= n2c32 (channels 1 .. of a k row) are stored in FP32, channel c = m / n2c32 as
// n2c32 float2 values in the first half of its FP64 slot (float2 index m + c n2c32 of the row).
// MT columns m per thread (m and m + 256, ...): every phase factor read from the LDS (a broadcast to the
// wave) serves MT outputs, which the many-output stages (P = mesh points of a line) are bound by.
#define WN_QMAX 16
#define WN_PCH 64
template <int MT, int QM>
__global__ __launch_bounds__(256) void k_contract(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                 const double2 *__restrict__ E, int P, int p0, int Q, int M,
                                                 long long ib, long long iq, long long ob, long long op, int n2c32) {
    extern __shared__ double2 es[];
    for (int e = threadIdx.x; e < P * Q; e += blockDim.x) es[e] = E[(size_t)p0 * Q + e];
    __syncthreads();
    const int mb = blockIdx.x * (256 * MT) + threadIdx.x;
    if (mb >= M) return;
    const long long b = blockIdx.y;
    double2 v[MT][QM];
    bool on[MT], f32[MT];
    long long dst[MT];
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        const int m = mb + 256 * u;
        on[u] = m < M;
        const double2 *src = in + b * ib + (on[u] ? m : 0);
#pragma unroll
        for (int q = 0; q < QM; ++q) v[u][q] = (on[u] && q < Q) ? src[q * iq] : make_double2(0.0, 0.0);
        f32[u] = n2c32 > 0 && m >= n2c32;
        // FP64: double2 index m of the row; FP32: float2 index m + (m / n2c32) n2c32
        dst[u] = f32[u] ? m + (long long)(m / n2c32) * n2c32 : m;
    }
    for (int p = 0; p < P; ++p) {
        const double2 *ep = es + p * Q;
        double2 acc[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) acc[u] = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < QM; ++q)
            if (q < Q) {
                const double2 e = ep[q];
#pragma unroll
                for (int u = 0; u < MT; ++u) acc[u] = cfma(e, v[u][q], acc[u]);
            }
        double2 *row = out + b * ob + p * op;
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            if (!on[u]) continue;
            if (f32[u])
                ((float2 *)row)[dst[u]] = make_float2((float)acc[u].x, (float)acc[u].y);
            else
                row[dst[u]] = acc[u];
        }
    }
}

// The same contraction for the stages with many outputs per input column (P >= 16: the points of a mesh line,
// the lines of a plane), as FP64 MFMA products: per input row b, out(P x M) = E(P x Q) V(Q x M) complex, with
// three real products (P1 = Er Vr, P2 = Ei Vi, P3 = (Er + Ei)(Vr + Vi); Re = P1 - P2, Im = P3 - P1 - P2) on
// 16x16x4 tiles.  Wave w of a workgroup owns output rows z 128 + 32 w .. + 31 (two 16-row tiles), whose operand
// slices of E stay in registers; the workgroup walks MC tiles of 16 columns m, each wave loading its 4 x 16
// slices of V from global memory one tile ahead.  No LDS, no barrier.
template <int QK, int MC>
__global__ __launch_bounds__(256) void k_contract_mfma(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                      const double2 *__restrict__ E, int P, int p0, int Q, int M,
                                                      long long ib, long long iq, long long ob, long long op,
                                                      int n2c32) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    const long long b = blockIdx.y;
    const int pbase = blockIdx.z * 128 + wave * 32;
    double er[2][QK], ei[2][QK];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < QK; ++s) {
            const int p = pbase + 16 * t + lr, q = 4 * s + lq;
            const double2 e = (p < P && q < Q) ? E[(size_t)(p0 + p) * Q + q] : make_double2(0.0, 0.0);
            er[t][s] = e.x;
            ei[t][s] = e.y;
        }
    const double2 *src = in + b * ib;
    auto loadv = [&](int mt, double2 (&vv)[QK]) {
        const int m = mt * 16 + lr;
#pragma unroll
        for (int s = 0; s < QK; ++s) {
            const int q = 4 * s + lq;
            vv[s] = (m < M && q < Q) ? src[q * iq + m] : make_double2(0.0, 0.0);
        }
    };
    const int mt0 = blockIdx.x * MC;
    double2 v[QK], nv[QK];
    loadv(mt0, v);
    for (int i = 0; i < MC; ++i) {
        const int mt = mt0 + i;
        if (mt * 16 >= M) break;
        if (i + 1 < MC) loadv(mt + 1, nv);
        const int m = mt * 16 + lr;
        const bool f32 = n2c32 > 0 && m >= n2c32;
        const long long dst = f32 ? m + (long long)(m / n2c32) * n2c32 : m;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (pbase + 16 * t >= P) break;
            f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < QK; ++s) {
                p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(er[t][s], v[s].x, p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ei[t][s], v[s].y, p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(er[t][s] + ei[t][s], v[s].x + v[s].y, p3, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = pbase + 16 * t + lq + 4 * j;
                if (p >= P || m >= M) continue;
                const double cr = p1[j] - p2[j], ci = p3[j] - (p1[j] + p2[j]);
                double2 *row = out + b * ob + (long long)p * op;
                // streaming stores: what one launch writes (a plane's line sums, 0.5 GB at n = 32, or its
                // channels, 3 GB) is read back only after more data than the caches hold
                if (f32) {
                    float *w = (float *)row + 2 * dst;
                    __builtin_nontemporal_store((float)cr, w);
                    __builtin_nontemporal_store((float)ci, w + 1);
                } else {
                    double *w = (double *)row + 2 * dst;
                    __builtin_nontemporal_store(cr, w);
                    __builtin_nontemporal_store(ci, w + 1);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < QK; ++s) v[s] = nv[s];
    }
}

hipError_t wn_launch_contract(const double2 *in, double2 *out, const double2 *E, int P, int p0, int Q, int M, int nb,
                              long long ib, long long iq, long long ob, long long op, hipStream_t st, int n2c32) {
    if (Q > WN_QMAX) return hipErrorInvalidValue;
    const size_t lds = (size_t)P * Q * sizeof(double2);
    // input terms held per column: the span Q rounded up to 8, 12 or 16
#define WN_CONTRACT(MT, QM)                                                                                      \
    hipLaunchKernelGGL((k_contract<MT, QM>), dim3((M + 256 * MT - 1) / (256 * MT), nb), dim3(256), lds, st, in, out, \
                       E, P, p0, Q, M, ib, iq, ob, op, n2c32)
#ifndef WN_CONTRACT_MFMA
#define WN_CONTRACT_MFMA 1
#endif
    if (WN_CONTRACT_MFMA && P >= 16) {
        constexpr int MC = 8;
        const dim3 grid((unsigned)((M + 16 * MC - 1) / (16 * MC)), (unsigned)nb, (unsigned)((P + 127) / 128));
        if (Q <= 8)
            hipLaunchKernelGGL((k_contract_mfma<2, MC>), grid, dim3(256), 0, st, in, out, E, P, p0, Q, M, ib, iq, ob,
                               op, n2c32);
        else if (Q <= 12)
            hipLaunchKernelGGL((k_contract_mfma<3, MC>), grid, dim3(256), 0, st, in, out, E, P, p0, Q, M, ib, iq, ob,
                               op, n2c32);
        else
            hipLaunchKernelGGL((k_contract_mfma<4, MC>), grid, dim3(256), 0, st, in, out, E, P, p0, Q, M, ib, iq, ob,
                               op, n2c32);
        return hipGetLastError();
    }
    if (P >= 8) {
        if (Q <= 8)
            WN_CONTRACT(2, 8);
        else if (Q <= 12)
            WN_CONTRACT(2, 12);
        else
            WN_CONTRACT(2, 16);
    } else {
        if (Q <= 8)
            WN_CONTRACT(1, 8);
        else if (Q <= 12)
            WN_CONTRACT(1, 12);
        else
            WN_CONTRACT(1, 16);
    }
#undef WN_CONTRACT
    return hipGetLastError();
}

// Selected outputs of the last contraction stage: group g = (b, t0, t1) writes out row t, t0 <= t < t1, as
// sum_{q<Q} E[sel[t] * Q + q] * in[b*ib + q*iq + m].  Each workgroup loads its Q input values per m once and
// produces the row's selected outputs (the representatives of one mesh row, written compactly).
__global__ __launch_bounds__(256) void k_contract_sel(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                     const double2 *__restrict__ E, const int4 *__restrict__ groups,
                                                     const int *__restrict__ sel, int Q, int M, long long ib,
                                                     long long iq, long long ot) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int4 gr = groups[blockIdx.y];
    const double2 *src = in + (long long)gr.x * ib + m;
    double2 v[WN_QMAX];
#pragma unroll
    for (int q = 0; q < WN_QMAX; ++q) v[q] = q < Q ? src[q * iq] : make_double2(0.0, 0.0);
    for (int t = gr.y; t < gr.z; ++t) {
        const double2 *ep = E + (size_t)sel[t] * Q;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < WN_QMAX; ++q)
            if (q < Q) acc = cfma(ep[q], v[q], acc);
        out[(long long)t * ot + m] = acc;
    }
}

hipError_t wn_launch_contract_sel(const double2 *in, double2 *out, const double2 *E, const int4 *groups, int ngroups,
                                  const int *sel, int Q, int M, long long ib, long long iq, long long ot,
                                  hipStream_t st) {
    if (Q > WN_QMAX) return hipErrorInvalidValue;
    if (ngroups <= 0) return hipSuccess;
    dim3 grid((M + 255) / 256, ngroups);
    hipLaunchKernelGGL(k_contract_sel, grid, dim3(256), 0, st, in, out, E, groups, sel, Q, M, ib, iq, ot);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Jacobi eigensolver: one workgroup (64, 128 or 256 threads) per k point, H and V in LDS (n <= 64).
//
// The upper triangle of channel 0 defines the Hermitian matrix (LAPACK uplo = 'U' semantics;
// the imaginary part of the diagonal is ignored).  Parallel cyclic Jacobi with a round-robin
// tournament: each round rotates n/2 disjoint (p,q) planes, first on columns (H J, V J), then on
// rows (J^+ H).  For h_pq = b e^{i phi}: J = diag(1, e^{-i phi}) [[c, s], [-s, c]] with
// t = sign(tau)/(|tau| + sqrt(1 + tau^2)), tau = (h_qq - h_pp)/(2b).  Sweeps stop when
// ||offdiag(H)||_F <= tol ||H||_F.  Eigenvalues are sorted ascending and every eigenvector
// column is rotated so that its first largest-magnitude entry is real and non-negative.
// ------------------------------------------------------------------------------------------------

// Block form of the parallel Jacobi round.  Only the upper triangle (and diagonal) of H is stored.
// A round with disjoint planes t = (p_t, q_t) maps every 2x2 block B = H[{p1,q1}][{p2,q2}] to
// J1^+ B J2, which reads and writes each stored element once; the diagonal blocks become
// diag(a - t b, d + t b) exactly.  V <- V J runs in the same phase.  Element (i, j) with i > j is
// read as conj(H[j][i]).
// Warm-started mesh eigensolver.  Consecutive k points (neighbours along the fastest mesh axis) form
// groups of WN_WARM_G; the anchor of a group (position WN_WARM_A, or the last point of a short group)
// is solved from the identity (pass 0).  Every other point k of the group is then solved from the
// anchor's eigenvectors U_a (pass 1): its channel 0 is first replaced by U_a^+ H(k) U_a, which is
// diagonal up to O(|k - k_a|), and V starts at U_a, so that V J1 J2 ... are eigenvectors of H(k).
// The stopping rule is unchanged; a nearby start only removes sweeps.
// pass < 0: every k point from the identity (grid = count); pass 0 / 1: the warm-start passes above;
// pass 2: every k point from uprev[k], the eigenvectors of the same point of the previous mesh plane
// (channel 0 already transformed by them).
// n > 64 runs the deferred-eigenvector solver of wnlqg_eigbig.hip instead (H and V exceed the LDS).
template <int N, int TPB>
__global__ __launch_bounds__(TPB, (N <= 32 && TPB == 256) ? 6 : 1) void k_eig_f64(const double *__restrict__ x, double *__restrict__ eout,
                                                double2 *__restrict__ uout, int *__restrict__ sweeps_out,
                                                unsigned long long *__restrict__ sweep_total, int n, int ldo,
                                                double tol, int pass, int count,
                                                const double2 *__restrict__ uprev, int ctr_stride,
                                                int *__restrict__ status, int max_sweeps) {
    constexpr int LD = N + 1;
    constexpr int HALF = N / 2;
    constexpr int VITEMS = N * HALF;
    constexpr int NW = TPB / 64;
    constexpr int NP = N * (N + 1) / 2;
    __shared__ double2 H[NP];       // packed upper triangle, row-major
    __shared__ double2 Vs[N * LD];
    __shared__ double4 rot[HALF];   // (c, s, cos phi, sin phi)
    __shared__ int2 pq[HALF];
    __shared__ double red[2 * NW];
    int anchor = -1;
    const int k = (pass < 0 || pass == 2) ? (int)blockIdx.x : wn_warm_k(blockIdx.x, pass, count, &anchor);
    if (k >= count) return;
    const int t = threadIdx.x;

    // packed row-major index of (i, j), i <= j (24-bit multiply: full-rate on the VALU)
    auto P = [](int i, int j) { return (int)(__umul24((unsigned)i, (unsigned)(2 * N - i + 1)) >> 1) + (j - i); };
    // V[i][j]: row i, column j (the eigenvector index)
    auto V = [&](int i, int j) -> double2 & { return Vs[i * LD + j]; };
    const double2 *xk = (const double2 *)(x + (size_t)k * ldo);
    const double2 *ua = pass == 2 ? uprev + (size_t)k * n * n : (anchor >= 0 ? uout + (size_t)anchor * n * n : nullptr);
    for (int e = t; e < N * N; e += TPB) {
        const int i = e / N, j = e - (e / N) * N;
        if (i <= j) {
            double2 hv = make_double2(0.0, 0.0);
            if (j < n) {
                hv = xk[i * n + j];
                if (i == j) hv.y = 0.0;
            }
            H[P(i, j)] = hv;
        }
        V(i, j) = (ua && i < n && j < n) ? ua[i * n + j] : make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();

    // Work split of a round: NOFF off-diagonal blocks (t1 < t2) and VITEMS updates of V.  When the
    // blocks fit in whole waves with at least one wave to spare (N = 32 at 256 threads), the first
    // BW threads take the blocks and the remaining waves take V, so that no wave carries both;
    // otherwise every thread takes its share of both.
    constexpr int NOFF = HALF * (HALF - 1) / 2;
    constexpr int BW0 = (NOFF + 63) / 64 * 64;
    constexpr bool SPLIT = BW0 + 64 <= TPB;
    constexpr int BW = SPLIT ? BW0 : TPB;           // threads updating blocks
    constexpr int VT0 = SPLIT ? BW0 : 0;            // first thread updating V
    constexpr int VW = SPLIT ? TPB - BW0 : TPB;     // threads updating V
    constexpr int BPT = (NOFF + BW - 1) / BW;
    constexpr int VPT = (VITEMS + VW - 1) / VW;     // V items per thread
    constexpr int VTPS = VW / HALF;                  // V threads per pair slot (slot-major mapping)
    constexpr bool VSLOT = VW % HALF == 0 && VTPS > 0 && N % VTPS == 0 && N / VTPS == VPT;
    // Bank-conflict-free V mapping: 16 consecutive lanes share one pair slot and take rows i0, i0 + 16, ...
    // (i0 = lane mod 16), so a 16-byte LDS access of the 16 lanes touches the 16 distinct 16-byte bank
    // groups (i0 (N + 1) + p) mod 16 (N + 1 odd); the VG16 lane groups take slots g, g + VG16, ...
    constexpr int VG16 = VW / 16;
    constexpr bool VQUAD = VW % 16 == 0 && N % 16 == 0 && VG16 > 0 && HALF % VG16 == 0 &&
                           (HALF / VG16) * (N / 16) == VPT && VT0 % 64 == 0;
    int bt1[BPT], bt2[BPT];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        // bidx -> (t1, t2), t1 < t2, row-major over the strict upper triangle of blocks
        const int bidx = t + q * BW;
        int t1 = 0, rem = bidx;
        while (t1 < HALF - 1 && rem >= HALF - 1 - t1) {
            rem -= HALF - 1 - t1;
            ++t1;
        }
        const bool ok = t < BW && bidx < NOFF;
        bt1[q] = ok ? t1 : -1;
        bt2[q] = ok ? t1 + 1 + rem : -1;
    }

    int sweep = 0;
    bool converged = false;
    double prev_off = 0.0;
    for (; sweep < max_sweeps; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int e = t; e < N * N; e += TPB) {
            const int i = e / N, j = e - (e / N) * N;
            if (i > j) continue;
            const double2 hv = H[P(i, j)];
            const double m2 = hv.x * hv.x + hv.y * hv.y;
            if (i == j)
                dia += m2;
            else
                off += 2.0 * m2;
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            dia += __shfl_xor(dia, o);
        }
        if (NW > 1) {
            if ((t & 63) == 0) {
                red[2 * (t >> 6)] = off;
                red[2 * (t >> 6) + 1] = dia;
            }
            __syncthreads();
            off = 0.0;
            dia = 0.0;
            for (int w = 0; w < NW; ++w) {
                off += red[2 * w];
                dia += red[2 * w + 1];
            }
            __syncthreads();
        }
        // converged, or at the rounding floor (a sweep no longer reduces the off-diagonal mass
        // quadratically once it is below ~1e-20 of the total)
        if (off <= tol * tol * (dia + off) ||
            (sweep > 0 && off > 0.25 * prev_off && off <= 1e-20 * (dia + off))) {
            converged = true;
            break;
        }
        prev_off = off;
        // round-robin slot t pairs p = (t - 1 + r) mod (N - 1) + 1 (0 for t = 0) with
        // q = (N - 2 - t + r) mod (N - 1) + 1: both residues advance by one per round
        int xp = t >= 1 && t < HALF ? t - 1 : 0, xq = t < HALF ? N - 2 - t : 0;
        for (int r = 0; r < N - 1; ++r) {
            if (t < HALF) {
                int p = (t == 0) ? 0 : xp + 1;
                int q = xq + 1;
                xp = xp == N - 2 ? 0 : xp + 1;
                xq = xq == N - 2 ? 0 : xq + 1;
                if (p > q) {
                    const int tmp = p;
                    p = q;
                    q = tmp;
                }
                double c = 1.0, s = 0.0, cph = 1.0, sph = 0.0;
                if (q < n) {
                    const double a = H[P(p, p)].x, d = H[P(q, q)].x;
                    const double2 h = H[P(p, q)];
                    const double b2 = h.x * h.x + h.y * h.y;
                    if (b2 > 0.0) {
                        // t = tan(theta) = sign(d-a) 2b / (|d-a| + sqrt((d-a)^2 + 4b^2)), the smaller root of
                        // t^2 + 2 tau t - 1 = 0 with tau = (d-a)/(2b)
                        const double ib = wn_rsqrt(b2);
                        const double b = b2 * ib;
                        const double diff = d - a;
                        const double rr = diff * diff + 4.0 * b2;
                        const double root = rr * wn_rsqrt(rr);
                        const double tt = (diff >= 0.0 ? 2.0 * b : -2.0 * b) * wn_rcp(fabs(diff) + root);
                        c = wn_rsqrt(1.0 + tt * tt);
                        s = tt * c;
                        cph = h.x * ib;
                        sph = h.y * ib;
                        H[P(p, p)] = make_double2(a - tt * b, 0.0);
                        H[P(q, q)] = make_double2(d + tt * b, 0.0);
                        H[P(p, q)] = make_double2(0.0, 0.0);
                    }
                }
                rot[t] = make_double4(c, s, cph, sph);
                pq[t] = make_int2(p, q);
            }
            __syncthreads();
            // off-diagonal blocks t1 < t2 of H
#pragma unroll
            for (int qb = 0; qb < BPT; ++qb) {
                const int t1 = bt1[qb], t2 = bt2[qb];
                if (t1 < 0) continue;
                const double4 R1 = rot[t1], R2 = rot[t2];
                if (R1.y == 0.0 && R2.y == 0.0) continue;
                const int2 a1 = pq[t1], a2 = pq[t2];
                const int rr[2] = {a1.x, a1.y}, cc[2] = {a2.x, a2.y};
                // element (i, j) is stored at (min, max) and conjugated when i > j: one 16-byte read each and
                // a sign flip of the imaginary part, without branches
                double2 B[2][2];
                int off[2][2];
                unsigned long long flip[2][2];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        const int i = rr[u], j = cc[v];
                        off[u][v] = P(min(i, j), max(i, j));
                        flip[u][v] = (unsigned long long)(i > j) << 63;
                    }
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        const double2 h = H[off[u][v]];
                        B[u][v] = make_double2(h.x, __longlong_as_double(__double_as_longlong(h.y) ^ flip[u][v]));
                    }
                // B J2: col0 = c2 B0 - s2 e^{-i phi2} B1, col1 = s2 B0 + c2 e^{-i phi2} B1
                const double2 em2 = make_double2(R2.z, -R2.w);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const double2 y = cmul(em2, B[u][1]);
                    const double2 x0 = B[u][0];
                    B[u][0] = make_double2(R2.x * x0.x - R2.y * y.x, R2.x * x0.y - R2.y * y.y);
                    B[u][1] = make_double2(R2.y * x0.x + R2.x * y.x, R2.y * x0.y + R2.x * y.y);
                }
                // J1^+ M: row0 = c1 M0 - s1 e^{i phi1} M1, row1 = s1 M0 + c1 e^{i phi1} M1
                const double2 ep1 = make_double2(R1.z, R1.w);
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const double2 y = cmul(ep1, B[1][v]);
                    const double2 x0 = B[0][v];
                    B[0][v] = make_double2(R1.x * x0.x - R1.y * y.x, R1.x * x0.y - R1.y * y.y);
                    B[1][v] = make_double2(R1.y * x0.x + R1.x * y.x, R1.y * x0.y + R1.x * y.y);
                }
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 2; ++v)
                        H[off[u][v]] = make_double2(B[u][v].x, __longlong_as_double(__double_as_longlong(B[u][v].y) ^
                                                                                   flip[u][v]));
            }
            // V <- V J.  Items (pair slot, row) are fixed per thread; a pair without rotation has
            // (c, s, cos phi, sin phi) = (1, 0, 1, 0), which maps the two entries to themselves exactly, so
            // every item is applied unconditionally; padded rows (i >= n) hold zeros in the columns of
            // rotated pairs (p, q < n) and stay zero.  When the slots divide the V threads, a thread keeps
            // one slot (one rotation read per round) and takes VPT rows of it.
            if (t >= VT0) {
                if constexpr (VQUAD) {
                    const int l = t - VT0, i0 = l & 15, g = l >> 4;
#pragma unroll
                    for (int sp = 0; sp < HALF / VG16; ++sp) {
                        const int pr = g + sp * VG16;
                        const double4 R = rot[pr];
                        const int2 ij = pq[pr];
#pragma unroll
                        for (int rq = 0; rq < N / 16; ++rq) {
                            const int i = i0 + rq * 16;
                            const double2 xp = V(i, ij.x), yq = V(i, ij.y);
                            const double2 ey = make_double2(R.z * yq.x + R.w * yq.y, R.z * yq.y - R.w * yq.x);
                            V(i, ij.x) = make_double2(R.x * xp.x - R.y * ey.x, R.x * xp.y - R.y * ey.y);
                            V(i, ij.y) = make_double2(R.y * xp.x + R.x * ey.x, R.y * xp.y + R.x * ey.y);
                        }
                    }
                } else if constexpr (VSLOT) {
                    const int pr = (t - VT0) / VTPS, i0 = (t - VT0) - pr * VTPS;
                    const double4 R = rot[pr];
                    const int2 ij = pq[pr];
#pragma unroll
                    for (int qv = 0; qv < VPT; ++qv) {
                        const int i = i0 + qv * VTPS;
                        const double2 xp = V(i, ij.x), yq = V(i, ij.y);
                        const double2 ey = make_double2(R.z * yq.x + R.w * yq.y, R.z * yq.y - R.w * yq.x);
                        V(i, ij.x) = make_double2(R.x * xp.x - R.y * ey.x, R.x * xp.y - R.y * ey.y);
                        V(i, ij.y) = make_double2(R.y * xp.x + R.x * ey.x, R.y * xp.y + R.x * ey.y);
                    }
                } else {
#pragma unroll
                    for (int qv = 0; qv < VPT; ++qv) {
                        const int e = t - VT0 + qv * VW;
                        if (e >= VITEMS) break;
                        const int pr = e / N, i = e - (e / N) * N;
                        const double4 R = rot[pr];
                        const int2 ij = pq[pr];
                        const double2 xp = V(i, ij.x), yq = V(i, ij.y);
                        const double2 ey = make_double2(R.z * yq.x + R.w * yq.y, R.z * yq.y - R.w * yq.x);
                        V(i, ij.x) = make_double2(R.x * xp.x - R.y * ey.x, R.x * xp.y - R.y * ey.y);
                        V(i, ij.y) = make_double2(R.y * xp.x + R.x * ey.x, R.y * xp.y + R.x * ey.y);
                    }
                }
            }
            __syncthreads();
        }
    }
    // sort ascending (stable by index); canonical phase per eigenvector column (first largest |v_i|
    // made real and non-negative), written in sorted order
    // rot / pq are free after the sweeps and H after its diagonal is read: ev, rank and phase reuse them
    double *ev = (double *)&rot[0];
    int *rank = (int *)&pq[0];
    double2 *phase = H;
    if (t < n) ev[t] = H[P(t, t)].x;
    __syncthreads();
    if (t < n) {
        const double et = ev[t];
        int rk = 0;
        for (int j = 0; j < n; ++j) {
            const double ej = ev[j];
            rk += (ej < et) || (ej == et && j < t);
        }
        rank[t] = rk;
        eout[(size_t)k * n + rk] = et;
        int piv = 0;
        double pm = hypot(V(0, t).x, V(0, t).y);
        for (int i = 1; i < n; ++i) {
            const double2 v = V(i, t);
            const double m = hypot(v.x, v.y);
            if (m > pm) {
                pm = m;
                piv = i;
            }
        }
        double2 ph = make_double2(1.0, 0.0);
        if (pm > 0.0) {
            const double2 v = V(piv, t);
            ph = make_double2(v.x / pm, -v.y / pm);
        }
        phase[t] = ph;
    }
    __syncthreads();
    double2 *uk = uout + (size_t)k * n * n;
    for (int e = t; e < n * n; e += TPB) {
        const int i = e / n, j = e - i * n;
        uk[i * n + rank[j]] = cmul(V(i, j), phase[j]);
    }
    // counters: bucket blockIdx % 64, ctr_stride words apart (0: one address)
    const int bucket = ctr_stride * (int)(blockIdx.x % WN_CTR_BUCKETS);
    if (t == 0 && sweeps_out) atomicMax(sweeps_out + bucket, sweep);
    if (t == 0 && sweep_total) atomicAdd(sweep_total + bucket / 2, (unsigned long long)sweep);
    // the sweep cap ended the solve before the stopping test passed: status word 1
    if (t == 0 && !converged && status) atomicOr(status + 1, 1);
}

hipError_t wn_launch_eig_pass_f64(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                                  int count, const WnShape &sh, int tpb, int pass, hipStream_t st, double2 *vscratch,
                                  const double2 *uprev, int ctr_stride, int *status) {
    if (sh.max_sweeps < 1) return hipErrorInvalidValue;
    if (pass == 2 && (!uprev || !wn_warm_start(sh.n, count))) return hipErrorInvalidValue;
    const int n = sh.n;
    if (wn_eig_rv_supported(n) && (tpb == 0 || tpb == WN_EIG_DV) && vscratch)
        return wn_launch_eig_defer(x, e, u, sweeps, sweep_total, count, sh, pass, st, uprev, ctr_stride, status,
                                   vscratch);
    if (wn_eig_rv_supported(n) && (tpb == 0 || tpb == WN_EIG_RV || tpb == WN_EIG_DV))
        return wn_launch_eig_rv(x, e, u, sweeps, sweep_total, count, sh, pass, st, uprev, ctr_stride, status);
    const double tol = 1.0e-15;
    const int groups = (count + WN_WARM_G - 1) / WN_WARM_G;
    const int grid = (pass < 0 || pass == 2) ? count : pass == 0 ? groups : groups * (WN_WARM_G - 1);
    // tpb: threads per k point (64, 128 or 256); 0 selects the default of the size class
#define WN_EIG_LAUNCH(NN, T) \
    hipLaunchKernelGGL((k_eig_f64<NN, T>), dim3(grid), dim3(T), 0, st, x, e, u, sweeps, sweep_total, n, sh.ldo, tol, \
                       pass, count, uprev, ctr_stride, status, sh.max_sweeps)
#define WN_EIG_CASE(NN, DEF)                   \
    if (n <= NN) {                             \
        const int t = tpb ? tpb : DEF;         \
        if (t == 64)                           \
            WN_EIG_LAUNCH(NN, 64);             \
        else if (t == 128)                     \
            WN_EIG_LAUNCH(NN, 128);            \
        else                                   \
            WN_EIG_LAUNCH(NN, 256);            \
        return hipGetLastError();              \
    }
    WN_EIG_CASE(8, 64)
    WN_EIG_CASE(16, 64)
    WN_EIG_CASE(24, 128)
    WN_EIG_CASE(32, 256)
    WN_EIG_CASE(48, 256)
    WN_EIG_CASE(64, 256)
#undef WN_EIG_CASE
#undef WN_EIG_LAUNCH
    // n in (64, 128]: deferred-eigenvector solver, every pass
    if (n <= 128) return wn_launch_eig_big(x, e, u, sweeps, sweep_total, count, sh, st, vscratch, ctr_stride, status,
                                           pass, uprev);
    return hipErrorInvalidValue;
}

// Capture-safe clears (no memset nodes in captured graphs).
__global__ void k_zero(double *__restrict__ p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_reset_flags(int *__restrict__ f) {
    for (int i = threadIdx.x; i < WN_FLAG_WORDS; i += blockDim.x) f[i] = i == 3 ? 0x7f7f7f7f : 0;
}

hipError_t wn_launch_zero(double *p, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t wn_launch_reset_flags(int *f, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(64), 0, st, f);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Hamiltonian-gauge rotation X_H = U^+ X U for channels 1..C-1 (in place), one 256-thread
// workgroup per k point; U and X staged in LDS, U^+ X kept in registers between the halves.
// ------------------------------------------------------------------------------------------------
#define WN_ROT_MAXO 16
__global__ __launch_bounds__(256) void k_rotate_f64(double *__restrict__ x, const double2 *__restrict__ u, int n,
                                                   int C, int ldo) {
    extern __shared__ double2 sm[];
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    const int ld = n + 1;
    const int n2 = n * n;
    double2 *Us = sm;
    double2 *Xs = Us + n * ld;
    const double2 *uk = u + (size_t)k * n2;
    for (int e = t; e < n2; e += 256) {
        int i = e / n, j = e - i * n;
        Us[i * ld + j] = uk[e];
    }
    double2 *xk = (double2 *)(x + (size_t)k * ldo);
    for (int c = 1; c < C; ++c) {
        double2 *xc = xk + (size_t)c * n2;
        __syncthreads();
        for (int e = t; e < n2; e += 256) {
            int i = e / n, j = e - i * n;
            Xs[i * ld + j] = xc[e];
        }
        __syncthreads();
        double2 tv[WN_ROT_MAXO];
#pragma unroll
        for (int o = 0; o < WN_ROT_MAXO; ++o) {
            const int e = t + 256 * o;
            if (e < n2) {
                int i = e / n, j = e - i * n;
                double2 acc = make_double2(0.0, 0.0);
                for (int l = 0; l < n; ++l) acc = cfma(cconj(Us[l * ld + i]), Xs[l * ld + j], acc);
                tv[o] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < WN_ROT_MAXO; ++o) {
            const int e = t + 256 * o;
            if (e < n2) {
                int i = e / n, j = e - i * n;
                Xs[i * ld + j] = tv[o];
            }
        }
        __syncthreads();
        for (int e = t; e < n2; e += 256) {
            int i = e / n, j = e - i * n;
            double2 acc = make_double2(0.0, 0.0);
            for (int l = 0; l < n; ++l) acc = cfma(Xs[i * ld + l], Us[l * ld + j], acc);
            xc[e] = acc;
        }
    }
}

// MFMA form of the rotation: complex products as three real products on v_mfma_f64_16x16x4f64
// (Gauss's three-multiplication form; the operand sums are formed in registers from the LDS reads),
//   T = U^+ X:  P1 = Ur^T Xr, P2 = Ui^T Xi, P3 = (Ur + Ui)^T (Xi - Xr);  Tr = P1 + P2, Ti = P3 + P1 - P2
//   Y = T U:    P1 = Tr Ur,   P2 = Ti Ui,   P3 = (Tr + Ti) (Ur + Ui);    Yr = P1 - P2, Yi = P3 - P1 - P2
// with U, X (then T) in LDS as planar real arrays padded to np = 16*ceil(n/16) (zeros beyond n).
// The (np/16)^2 output tiles are dealt to the 4 waves.
// WARM: the warm-start transform of the mesh eigensolver: channel 0's upper triangle replaced by that of W^+ H W
// (H the Hermitian matrix of channel 0's upper triangle, read in row order), with W the
// group anchor's eigenvectors for the pass-1 points (uw = nullptr), or W = uw[k], the eigenvectors of
// the same point of the previous mesh plane, for every point (uw set).  Otherwise channels 1 .. C-1 of
// every k point with its own U.
template <int NP, int PF, bool WARM>
__global__ __launch_bounds__(256) void k_rotate_mfma(double *__restrict__ x, const double2 *__restrict__ u, int n,
                                                    int C, int ldo, int count, const double2 *__restrict__ uw) {
    extern __shared__ double smd[];
    int ka = blockIdx.x;
    const int k = (WARM && !uw) ? wn_warm_k(blockIdx.x, 1, count, &ka) : (int)blockIdx.x;
    if (WARM && k >= count) return;
    if (WARM && uw) u = uw;
    const int cbeg = WARM ? 0 : 1, cend = WARM ? 1 : C;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    constexpr int np = NP;     // n rounded up to 16 (compile time: the k loops unroll)
    constexpr int ld = np + 1;
    constexpr int nt = np >> 4;
    double *Ur = smd;
    double *Ui = Ur + np * ld;
    double *Xr = Ui + np * ld;
    double *Xi = Xr + np * ld;
    const double2 *uk = u + (size_t)ka * n * n;
    for (int e = t; e < np * np; e += 256) {
        int i = e / np, j = e - i * np;
        double2 v = (i < n && j < n) ? uk[i * n + j] : make_double2(0.0, 0.0);
        Ur[i * ld + j] = v.x;
        Ui[i * ld + j] = v.y;
    }
    double2 *xk = (double2 *)(x + (size_t)k * ldo);
    const int lr = lane & 15, lq = lane >> 4;
    // channel c + 1 is loaded into registers while channel c's second product runs (PF = np^2 / 256)
    double2 pf[PF];
    auto prefetch = [&](int c) {
        const double2 *xc = xk + (size_t)c * n * n;
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int e = t + 256 * q;
            const int i = e / np, j = e - (e / np) * np;
            // (WARM: every element read in row order; the Hermitian matrix of the upper triangle, the eigensolver's
            // input, is formed when the tile is stored)
            pf[q] = (e < np * np && i < n && j < n) ? xc[i * n + j] : make_double2(0.0, 0.0);
        }
    };
    if (cbeg < cend) prefetch(cbeg);
    for (int c = cbeg; c < cend; ++c) {
        double2 *xc = xk + (size_t)c * n * n;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int e = t + 256 * q;
            if (e < np * np) {
                const int i = e / np, j = e - (e / np) * np;
                if (!WARM) {
                    Xr[i * ld + j] = pf[q].x;
                    Xi[i * ld + j] = pf[q].y;
                } else if (j >= i) {
                    // upper-triangle element (i, j) and its conjugate at (j, i); the diagonal is real
                    Xr[i * ld + j] = pf[q].x;
                    Xi[i * ld + j] = i == j ? 0.0 : pf[q].y;
                    Xr[j * ld + i] = pf[q].x;
                    Xi[j * ld + i] = i == j ? 0.0 : -pf[q].y;
                }
            }
        }
        __syncthreads();
        // T = U^+ X, tiles kept in registers (at most 4 per wave for np <= 64)
        f64x4 tr[4], ti[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int tile = wave + 4 * s4;
            f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
            if (tile < nt * nt) {
                const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
#pragma unroll
                for (int k0 = 0; k0 < np; k0 += 4) {
                    const int kr = k0 + lq;
                    const double ur = Ur[kr * ld + r0 + lr], ui = Ui[kr * ld + r0 + lr];
                    const double xr = Xr[kr * ld + c0 + lr], xi = Xi[kr * ld + c0 + lr];
                    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur, xr, p1, 0, 0, 0);
                    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ui, xi, p2, 0, 0, 0);
                    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur + ui, xi - xr, p3, 0, 0, 0);
                }
            }
            // Tr = P1 + P2, Ti = P3 + P1 - P2
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                tr[s4][q] = p1[q] + p2[q];
                ti[s4][q] = p3[q] + (p1[q] - p2[q]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int tile = wave + 4 * s4;
            if (tile < nt * nt) {
                const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Xr[(r0 + lq + 4 * q) * ld + c0 + lr] = tr[s4][q];
                    Xi[(r0 + lq + 4 * q) * ld + c0 + lr] = ti[s4][q];
                }
            }
        }
        if (c + 1 < cend) prefetch(c + 1);
        __syncthreads();
        // Y = T U, written to global (WARM: the upper triangle only, which is what the eigensolver reads)
        for (int tile = wave; tile < nt * nt; tile += 4) {
            const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
            if (WARM && r0 > c0) continue;   // a tile below the diagonal
            f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < np; k0 += 4) {
                const int kc = k0 + lq;
                const double tr_ = Xr[(r0 + lr) * ld + kc], ti_ = Xi[(r0 + lr) * ld + kc];
                const double ur = Ur[kc * ld + c0 + lr], ui = Ui[kc * ld + c0 + lr];
                p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_, ur, p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti_, ui, p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_ + ti_, ur + ui, p3, 0, 0, 0);
            }
            // Yr = P1 - P2, Yi = P3 - P1 - P2
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = r0 + lq + 4 * q, col = c0 + lr;
                if (row < n && col < n && (!WARM || col >= row))
                    xc[row * n + col] = make_double2(p1[q] - p2[q], p3[q] - (p1[q] + p2[q]));
            }
        }
    }
}

// Rotation for the shift/injection-current pair kernels (16 <= n <= 64, FP64).  A point without active
// pairs is skipped (nothing reads its channels).  A point with at most WN_PCAP active pairs rotates the
// velocity and position channels (dH_a, A_a: the pair kernels sum their rows and columns over all
// intermediate bands) in full, and of dd H and dA only the matrix elements the pairs use: with B the bands
// of the pairs (at most 16), T_B = U_B^+ X (one 16-row MFMA tile per column tile), then
// (U^+ X U)_ab = sum_q T_B[a][q] U[q][b] for (a, b) = (n, m) and (m, n), written to
// pwd[k][pair][direction][15] (ddH channels 0..5, then dA 6..14) and not back to the channel.  Points with
// more pairs rotate every channel.  Products as in k_rotate_mfma (three real MFMA products per complex
// product).
// C32: channels 1 .. C-1 arrive in FP32 (the first half of each channel's FP64 slot, n^2 float2 values; the
// mixed form's mesh channels); the products are FP64 and every channel written back is FP64.
template <int NP, int PF, bool C32>
#ifndef WN_ROTSC_WAVES
#define WN_ROTSC_WAVES 3     // FP32 channels
#endif
#ifndef WN_ROTSC_WAVES64
#define WN_ROTSC_WAVES64 4   // FP64 channels
#endif
#ifndef WN_ROTSC_REG64
// register-operand loop for FP64 channels too: measured slower there (its double-width prefetch spills at
// four waves per SIMD, 283 ms per 128^3 mesh, and runs at 215 ms with three) than the staged loop (193 ms)
#define WN_ROTSC_REG64 0
#endif
// np = 32 (34 KB of LDS per workgroup): registers for three waves per SIMD with FP32 channels (the register
// operand loop carries two channels' X slices and the next two), four with FP64 channels
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256),
                          amdgpu_waves_per_eu(NP == 32 ? (C32 ? WN_ROTSC_WAVES : WN_ROTSC_WAVES64) : 1))) void
k_rotate_sc(double *__restrict__ x, const double2 *__restrict__ u,
                                                  const int *__restrict__ count, const int4 *__restrict__ pmeta,
                                                  double2 *__restrict__ pwd, int n, int C, int ldo,
                                                  int *__restrict__ flags, int herm) {
    extern __shared__ double smd[];
    __shared__ int bl[16];            // bands of the pairs
    __shared__ int prow[WN_PCAP][2];  // row of n and of m in bl, per pair
    __shared__ int pband[WN_PCAP][2]; // n and m
    __shared__ int nbl;
    const int k = blockIdx.x;
    const int cnt = count[k];
    if (cnt == 0) return;
    const bool sparse = cnt <= WN_PCAP;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    constexpr int np = NP;
    constexpr int ld = np + 1;
    constexpr int nt = np >> 4;
    double *Ur = smd;
    double *Ui = Ur + np * ld;
    double *Xr = Ui + np * ld;
    double *Xi = Xr + np * ld;
    const double2 *uk = u + (size_t)k * n * n;
    if (sparse && t == 0) {
        int nb_ = 0;
        for (int p = 0; p < cnt; ++p) {
            const int4 mt = pmeta[(size_t)k * n * n + p];
            const int ab[2] = {mt.x, mt.y};
            for (int d = 0; d < 2; ++d) {
                int r = 0;
                while (r < nb_ && bl[r] != ab[d]) ++r;
                if (r == nb_) bl[nb_++] = ab[d];
                prow[p][d] = r;
                pband[p][d] = ab[d];
            }
        }
        nbl = nb_;
    }
    if (t == 0) {
        int *bk = flags + WN_CTR_BASE + WN_CTR_STRIDE * (k % WN_CTR_BUCKETS);
        if (sparse) {
            atomicAdd(bk + 6, 1);
            atomicAdd(bk + 8, nbl);
            if (herm) atomicAdd(bk + 9, 1);
        } else {
            atomicAdd(bk + 7, 1);
        }
    }
    for (int e = t; e < np * np; e += 256) {
        const int i = e / np, j = e - i * np;
        const double2 v = (i < n && j < n) ? uk[i * n + j] : make_double2(0.0, 0.0);
        Ur[i * ld + j] = v.x;
        Ui[i * ld + j] = v.y;
    }
    double2 *xk = (double2 *)(x + (size_t)k * ldo);
    const int lr = lane & 15, lq = lane >> 4;
    // K splits of the 16-row products (np = 32: two column tiles on four waves)
    constexpr int KS = np == 32 ? 2 : 1;
    if (sparse && herm && (C32 || WN_ROTSC_REG64)) {
        // Hermitian model, at most WN_PCAP pairs: every channel's product is U_B^+ X (rows at the pairs' bands) and,
        // for dH and A, (U_B^+ X) U.  Wave w owns column tile w % nt and K range w / nt of both products, so its
        // MFMA operand slices of U_B^+ and U are the same for all 21 channels and stay in registers, and its slice
        // of X (16 columns x np / KS rows) is loaded from global memory into registers one channel ahead; only
        // the partial results go through the LDS.  Up to 8 bands the rows are one or two 4-row tiles on the
        // four-block 4 x 4 x 4 MFMA (about 20 cycles against 64 for a 16-row tile): block b holds columns 4b ..
        // 4b + 3 of the wave's 16, lane 16k + 4b + i carries A[i][k], lane 16k + 4b + j carries B[k][j] and
        // element (i, j) comes back in lane 16i + 4b + j, so lane (lq, lr) holds row 4 t + lq of tile t at column
        // c0 + lr, and the X slice is the 16 x 16 x 4 form's.  More bands: one 16-row tile (rows lq + 4 q).
        constexpr int KST = np / KS / 4;   // MFMA K steps of a wave
        const int tile = wave % nt, kp = wave / nt;
        const bool tw = wave < nt * KS;
        const int c0 = tile * 16, kb = kp * (np / KS);
        __syncthreads();   // U staged
        const int nb_ = nbl;
        double bur[KST], bui[KST];
#pragma unroll
        for (int s = 0; s < KST; ++s) {
            const int kr = kb + 4 * s + lq;
            bur[s] = Ur[kr * ld + c0 + lr];
            bui[s] = Ui[kr * ld + c0 + lr];
        }
        const bool colok = tw && c0 + lr < n;
        // X[kr][c0 + lr] of channel c for this wave's K steps, kept in the channel's storage precision until used
        using XT = typename std::conditional<C32, float2, double2>::type;
        auto load_x = [&](int c, XT (&xv)[KST]) {
            const XT *xc = (const XT *)(xk + (size_t)c * n * n);
#pragma unroll
            for (int s = 0; s < KST; ++s) {
                const int kr = kb + 4 * s + lq;
                xv[s] = (colok && kr < n) ? xc[kr * n + c0 + lr] : XT{};
            }
        };
        // more than 8 bands: one 16-row tile per channel (rows lq + 4 q of the result in slot q of a lane)
        auto channels16 = [&]() {
            constexpr int NQ = 4;
            auto rrow = [&](int q) { return lq + 4 * q; };
            // A operand slices of U_B^+ (row lr of the tile)
            double aur[KST], aui[KST];
#pragma unroll
            for (int s = 0; s < KST; ++s) {
                const int kr = kb + 4 * s + lq;
                const int br = lr < nb_ ? bl[lr] : 0;
                const double rm = lr < nb_ ? 1.0 : 0.0;
                aur[s] = rm * Ur[kr * ld + br];
                aui[s] = rm * Ui[kr * ld + br];
            }
            XT xv[KST], nxv[KST];
            load_x(1, xv);
            for (int c = 1; c < C; ++c) {
                double2 *xc = xk + (size_t)c * n * n;
                const bool full = (c >= WN_CH_DH && c < WN_CH_DDH) || (c >= WN_CH_A && c < WN_CH_DA);
                if (c + 1 < C) load_x(c + 1, nxv);
                double xr[KST], xi[KST];
#pragma unroll
                for (int s = 0; s < KST; ++s) {
                    xr[s] = (double)xv[s].x;
                    xi[s] = (double)xv[s].y;
                }
                // partial T_B = U_B^+ X over this wave's K range
                double tr[NQ], ti[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) tr[q] = ti[q] = 0.0;
                if (tw) {
                    f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s = 0; s < KST; ++s) {
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(aur[s], xr[s], p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(aui[s], xi[s], p2, 0, 0, 0);
                        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(aur[s] + aui[s], xi[s] - xr[s], p3, 0, 0,
                                                                  0);
                    }
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        tr[q] = p1[q] + p2[q];
                        ti[q] = p3[q] + (p1[q] - p2[q]);
                    }
                }
                __syncthreads();   // the previous channel's readers of the partial rows are done
                if (tw) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        Xr[(16 * kp + rrow(q)) * ld + c0 + lr] = tr[q];
                        Xi[(16 * kp + rrow(q)) * ld + c0 + lr] = ti[q];
                    }
                }
                __syncthreads();
                if (full) {
                    // rows R_B = T_B U over this wave's K range, written at rows bl[r] (the pair kernel reads the
                    // other entries of the pairs' columns as conjugates of these rows)
                    double yr[NQ], yi[NQ];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) yr[q] = yi[q] = 0.0;
                    if (tw) {
                        f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int s = 0; s < KST; ++s) {
                            const int kc = kb + 4 * s + lq;
                            double tr_ = Xr[lr * ld + kc], ti_ = Xi[lr * ld + kc];
                            if constexpr (KS == 2) {
                                tr_ += Xr[(16 + lr) * ld + kc];
                                ti_ += Xi[(16 + lr) * ld + kc];
                            }
                            p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_, bur[s], p1, 0, 0, 0);
                            p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti_, bui[s], p2, 0, 0, 0);
                            p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_ + ti_, bur[s] + bui[s], p3, 0, 0, 0);
                        }
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            yr[q] = p1[q] - p2[q];
                            yi[q] = p3[q] - (p1[q] + p2[q]);
                        }
                    }
                    if constexpr (KS == 2) {
                        __syncthreads();
                        if (tw && kp == 1) {
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                Xr[rrow(q) * ld + c0 + lr] = yr[q];
                                Xi[rrow(q) * ld + c0 + lr] = yi[q];
                            }
                        }
                        __syncthreads();
                    }
                    if (tw && kp == 0) {
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const int r = rrow(q), col = c0 + lr;
                            double vr = yr[q], vi = yi[q];
                            if constexpr (KS == 2) {
                                vr += Xr[r * ld + col];
                                vi += Xi[r * ld + col];
                            }
                            if (r < nb_ && col < n) xc[bl[r] * n + col] = make_double2(vr, vi);
                        }
                    }
                } else {
                    // the pairs' elements sum_q T_B[a][q] U[q][b]: 16 lanes per (pair, direction), butterfly sum
                    const int item = t >> 4, sub = t & 15;
                    const int p = item >> 1, d = item & 1;
                    const bool ok = p < cnt;
                    const int ra = ok ? prow[p][d] : 0, bb = ok ? pband[p][d ^ 1] : 0;
                    double sr = 0.0, si = 0.0;
                    if (ok) {
                        for (int q = sub; q < np; q += 16) {
                            double tr_ = Xr[ra * ld + q], ti_ = Xi[ra * ld + q];
                            if constexpr (KS == 2) {
                                tr_ += Xr[(16 + ra) * ld + q];
                                ti_ += Xi[(16 + ra) * ld + q];
                            }
                            const double ur = Ur[q * ld + bb], ui = Ui[q * ld + bb];
                            sr += tr_ * ur - ti_ * ui;
                            si += tr_ * ui + ti_ * ur;
                        }
                    }
                    for (int o = 8; o > 0; o >>= 1) {
                        sr += __shfl_xor(sr, o);
                        si += __shfl_xor(si, o);
                    }
                    if (ok && sub == 0) {
                        const int slot = c < WN_CH_A ? c - WN_CH_DDH : 6 + (c - WN_CH_DA);
                        pwd[(((size_t)k * WN_PCAP + p) * 2 + d) * 15 + slot] = make_double2(sr, si);
                    }
                }
#pragma unroll
                for (int s = 0; s < KST; ++s) xv[s] = nxv[s];
            }
        };
        // Points with <= 8 bands, two channels per barrier phase: the 4-row tiles' partial rows of channel c sit
        // at rows 16 kp + r and those of channel c + 1 at rows 8 + 16 kp + r (r < 8), so one pair of barriers
        // serves both channels (and one more pair the K-split exchange of the full channels among them).
        auto channel_pairs = [&]() {
            double aur[2][KST], aui[2][KST];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int s = 0; s < KST; ++s) {
                    const int kr = kb + 4 * s + lq;
                    const int row = 4 * rt + (lane & 3);
                    const int br = row < nb_ ? bl[row] : 0;
                    const double rm = row < nb_ ? 1.0 : 0.0;
                    aur[rt][s] = rm * Ur[kr * ld + br];
                    aui[rt][s] = rm * Ui[kr * ld + br];
                }
            auto is_full = [](int c) { return (c >= WN_CH_DH && c < WN_CH_DDH) || (c >= WN_CH_A && c < WN_CH_DA); };
            XT xv[2][KST], nxv[2][KST];
            load_x(1, xv[0]);
            if (2 < C) load_x(2, xv[1]);
            for (int c = 1; c < C; c += 2) {
                const bool hasb = c + 1 < C;
                if (c + 2 < C) load_x(c + 2, nxv[0]);
                if (c + 3 < C) load_x(c + 3, nxv[1]);
                // partial T_B = U_B^+ X over this wave's K range, both channels (rows 4 rt + lq, column c0 + lr)
                double tr[2][2], ti[2][2];
#pragma unroll
                for (int ch = 0; ch < 2; ++ch) {
                    tr[ch][0] = tr[ch][1] = ti[ch][0] = ti[ch][1] = 0.0;
                    if (!tw || (ch == 1 && !hasb)) continue;
                    double xr[KST], xi[KST];
#pragma unroll
                    for (int s = 0; s < KST; ++s) {
                        xr[s] = (double)xv[ch][s].x;
                        xi[s] = (double)xv[ch][s].y;
                    }
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        if (rt * 4 >= nb_) break;   // a tile of padding rows only (zero)
                        double p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
                        for (int s = 0; s < KST; ++s) {
                            p1 = __builtin_amdgcn_mfma_f64_4x4x4f64(aur[rt][s], xr[s], p1, 0, 0, 0);
                            p2 = __builtin_amdgcn_mfma_f64_4x4x4f64(aui[rt][s], xi[s], p2, 0, 0, 0);
                            p3 = __builtin_amdgcn_mfma_f64_4x4x4f64(aur[rt][s] + aui[rt][s], xi[s] - xr[s], p3, 0,
                                                                    0, 0);
                        }
                        tr[ch][rt] = p1 + p2;
                        ti[ch][rt] = p3 + (p1 - p2);
                    }
                }
                __syncthreads();   // the previous pair's readers of the partial rows are done
                if (tw) {
#pragma unroll
                    for (int ch = 0; ch < 2; ++ch)
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            Xr[(8 * ch + 16 * kp + 4 * q + lq) * ld + c0 + lr] = tr[ch][q];
                            Xi[(8 * ch + 16 * kp + 4 * q + lq) * ld + c0 + lr] = ti[ch][q];
                        }
                }
                __syncthreads();
                const bool fl[2] = {is_full(c), hasb && is_full(c + 1)};
                double yr[2][2], yi[2][2];
#pragma unroll
                for (int ch = 0; ch < 2; ++ch) {
                    yr[ch][0] = yr[ch][1] = yi[ch][0] = yi[ch][1] = 0.0;
                    if (ch == 1 && !hasb) continue;
                    if (fl[ch]) {
                        // rows R_B = T_B U over this wave's K range
                        if (!tw) continue;
#pragma unroll
                        for (int rt = 0; rt < 2; ++rt) {
                            if (rt * 4 >= nb_) break;
                            double p1 = 0.0, p2 = 0.0, p3 = 0.0;
                            const int row = 8 * ch + 4 * rt + (lane & 3);
#pragma unroll
                            for (int s = 0; s < KST; ++s) {
                                const int kc = kb + 4 * s + lq;
                                double tr_ = Xr[row * ld + kc], ti_ = Xi[row * ld + kc];
                                if constexpr (KS == 2) {
                                    tr_ += Xr[(16 + row) * ld + kc];
                                    ti_ += Xi[(16 + row) * ld + kc];
                                }
                                p1 = __builtin_amdgcn_mfma_f64_4x4x4f64(tr_, bur[s], p1, 0, 0, 0);
                                p2 = __builtin_amdgcn_mfma_f64_4x4x4f64(ti_, bui[s], p2, 0, 0, 0);
                                p3 = __builtin_amdgcn_mfma_f64_4x4x4f64(tr_ + ti_, bur[s] + bui[s], p3, 0, 0, 0);
                            }
                            yr[ch][rt] = p1 - p2;
                            yi[ch][rt] = p3 - (p1 + p2);
                        }
                    } else {
                        // the pairs' elements sum_q T_B[a][q] U[q][b]: 16 lanes per (pair, direction)
                        const int item = t >> 4, sub = t & 15;
                        const int p = item >> 1, d = item & 1;
                        const bool ok = p < cnt;
                        const int ra = ok ? 8 * ch + prow[p][d] : 0, bb = ok ? pband[p][d ^ 1] : 0;
                        double sr = 0.0, si = 0.0;
                        if (ok) {
                            for (int q = sub; q < np; q += 16) {
                                double tr_ = Xr[ra * ld + q], ti_ = Xi[ra * ld + q];
                                if constexpr (KS == 2) {
                                    tr_ += Xr[(16 + ra) * ld + q];
                                    ti_ += Xi[(16 + ra) * ld + q];
                                }
                                const double ur = Ur[q * ld + bb], ui = Ui[q * ld + bb];
                                sr += tr_ * ur - ti_ * ui;
                                si += tr_ * ui + ti_ * ur;
                            }
                        }
                        for (int o = 8; o > 0; o >>= 1) {
                            sr += __shfl_xor(sr, o);
                            si += __shfl_xor(si, o);
                        }
                        if (ok && sub == 0) {
                            const int cc = c + ch;
                            const int slot = cc < WN_CH_A ? cc - WN_CH_DDH : 6 + (cc - WN_CH_DA);
                            pwd[(((size_t)k * WN_PCAP + p) * 2 + d) * 15 + slot] = make_double2(sr, si);
                        }
                    }
                }
                if (fl[0] || fl[1]) {
                    if constexpr (KS == 2) {
                        __syncthreads();   // every read of the partial rows is done
                        if (tw && kp == 1) {
#pragma unroll
                            for (int ch = 0; ch < 2; ++ch) {
                                if (!fl[ch]) continue;
#pragma unroll
                                for (int q = 0; q < 2; ++q) {
                                    Xr[(8 * ch + 4 * q + lq) * ld + c0 + lr] = yr[ch][q];
                                    Xi[(8 * ch + 4 * q + lq) * ld + c0 + lr] = yi[ch][q];
                                }
                            }
                        }
                        __syncthreads();
                    }
                    if (tw && kp == 0) {
#pragma unroll
                        for (int ch = 0; ch < 2; ++ch) {
                            if (!fl[ch]) continue;
                            double2 *xc = xk + (size_t)(c + ch) * n * n;
#pragma unroll
                            for (int q = 0; q < 2; ++q) {
                                const int r = 4 * q + lq, col = c0 + lr;
                                double vr = yr[ch][q], vi = yi[ch][q];
                                if constexpr (KS == 2) {
                                    vr += Xr[(8 * ch + r) * ld + col];
                                    vi += Xi[(8 * ch + r) * ld + col];
                                }
                                if (r < nb_ && col < n) xc[bl[r] * n + col] = make_double2(vr, vi);
                            }
                        }
                    }
                }
#pragma unroll
                for (int ch = 0; ch < 2; ++ch)
#pragma unroll
                    for (int s = 0; s < KST; ++s) xv[ch][s] = nxv[ch][s];
            }
        };
        if (nb_ <= 8)
            channel_pairs();
        else
            channels16();
        return;
    }
    // FP32 channels (n even): pairs of consecutive elements per 16-byte load when the tile splits into pairs
    // per thread
    constexpr bool C32P = PF % 2 == 0;
    double2 pf[PF];
    auto prefetch = [&](int c) {
        const double2 *xc = xk + (size_t)c * n * n;
        const float2 *xc32 = (const float2 *)xc;
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int e = t + 256 * q;
            const int i = e / np, j = e - (e / np) * np;
            if constexpr (C32 && C32P) {
                // two consecutive elements (e0, e0 + 1 of the row-major tile) per 16-byte load; n is even, so
                // every pair lies in one row and is 16-byte aligned
                if ((q & 1) == 0) {
                    const int e0 = 2 * t + 512 * (q >> 1);
                    const int i0 = e0 / np, j0 = e0 - (e0 / np) * np;
                    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (e0 < np * np && i0 < n && j0 < n) v = *(const float4 *)(xc32 + i0 * n + j0);
                    pf[q] = make_double2((double)v.x, (double)v.y);
                    pf[q + 1] = make_double2((double)v.z, (double)v.w);
                }
            } else if constexpr (C32) {
                const float2 v = (e < np * np && i < n && j < n) ? xc32[i * n + j] : make_float2(0.0f, 0.0f);
                pf[q] = make_double2((double)v.x, (double)v.y);
            } else {
                pf[q] = (e < np * np && i < n && j < n) ? xc[i * n + j] : make_double2(0.0, 0.0);
            }
        }
    };
    prefetch(1);
    for (int c = 1; c < C; ++c) {
        double2 *xc = xk + (size_t)c * n * n;
        const bool full = !sparse || (c >= WN_CH_DH && c < WN_CH_DDH) || (c >= WN_CH_A && c < WN_CH_DA);
        // Hermitian model, few pairs: the velocity and position channels as rows at the pairs' bands
        const bool rows = full && sparse && herm;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int e = C32 && C32P ? 2 * t + 512 * (q >> 1) + (q & 1) : t + 256 * q;
            if (e < np * np) {
                const int i = e / np, j = e - (e / np) * np;
                Xr[i * ld + j] = pf[q].x;
                Xi[i * ld + j] = pf[q].y;
            }
        }
        // the next channel's loads are in flight during this channel's products
        if (c + 1 < C) prefetch(c + 1);
        __syncthreads();
        if (full && !rows) {
            // T = U^+ X in registers, then to LDS
            f64x4 tr[4], ti[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int tile = wave + 4 * s4;
                f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
                if (tile < nt * nt) {
                    const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
#pragma unroll
                    for (int k0 = 0; k0 < np; k0 += 4) {
                        const int kr = k0 + lq;
                        const double ur = Ur[kr * ld + r0 + lr], ui = Ui[kr * ld + r0 + lr];
                        const double xr = Xr[kr * ld + c0 + lr], xi = Xi[kr * ld + c0 + lr];
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur, xr, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ui, xi, p2, 0, 0, 0);
                        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur + ui, xi - xr, p3, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    tr[s4][q] = p1[q] + p2[q];
                    ti[s4][q] = p3[q] + (p1[q] - p2[q]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int tile = wave + 4 * s4;
                if (tile < nt * nt) {
                    const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        Xr[(r0 + lq + 4 * q) * ld + c0 + lr] = tr[s4][q];
                        Xi[(r0 + lq + 4 * q) * ld + c0 + lr] = ti[s4][q];
                    }
                }
            }
            __syncthreads();
            // Y = T U, written to global
            for (int tile = wave; tile < nt * nt; tile += 4) {
                const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
                f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = 0; k0 < np; k0 += 4) {
                    const int kc = k0 + lq;
                    const double tr_ = Xr[(r0 + lr) * ld + kc], ti_ = Xi[(r0 + lr) * ld + kc];
                    const double ur = Ur[kc * ld + c0 + lr], ui = Ui[kc * ld + c0 + lr];
                    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_, ur, p1, 0, 0, 0);
                    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti_, ui, p2, 0, 0, 0);
                    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_ + ti_, ur + ui, p3, 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + lq + 4 * q, col = c0 + lr;
                    if (row < n && col < n) xc[row * n + col] = make_double2(p1[q] - p2[q], p3[q] - (p1[q] + p2[q]));
                }
            }
        } else {
            // T_B = U_B^+ X: one 16-row tile (row r <-> band bl[r], zero rows beyond nbl) by nt column tiles; at
            // np = 32 the two column tiles' K range is split over the wave pairs (every wave busy), the partial
            // tiles landing in rows 16 kp .. 16 kp + 15 of the X area and summed where they are read
            const int nb_ = nbl;
            const int kp = wave / nt;
            const bool tw = wave < nt * KS;
            f64x4 tr = {0.0, 0.0, 0.0, 0.0}, ti = {0.0, 0.0, 0.0, 0.0};
            if (tw) {
                const int c0 = (wave % nt) * 16;
                const int br = lr < nb_ ? bl[lr] : 0;
                const double rm = lr < nb_ ? 1.0 : 0.0;
                f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = kp * (np / KS); k0 < (kp + 1) * (np / KS); k0 += 4) {
                    const int kr = k0 + lq;
                    const double ur = rm * Ur[kr * ld + br], ui = rm * Ui[kr * ld + br];
                    const double xr = Xr[kr * ld + c0 + lr], xi = Xi[kr * ld + c0 + lr];
                    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur, xr, p1, 0, 0, 0);
                    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ui, xi, p2, 0, 0, 0);
                    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ur + ui, xi - xr, p3, 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    tr[q] = p1[q] + p2[q];
                    ti[q] = p3[q] + (p1[q] - p2[q]);
                }
            }
            __syncthreads();
            if (tw) {
                const int c0 = (wave % nt) * 16;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Xr[(16 * kp + lq + 4 * q) * ld + c0 + lr] = tr[q];
                    Xi[(16 * kp + lq + 4 * q) * ld + c0 + lr] = ti[q];
                }
            }
            __syncthreads();
            if (rows) {
                // rows of the Hermitian U^+ X U at the bands B: R_B = T_B U (K split as above), written at rows
                // bl[r] (the pair kernel reads the other entries of the pairs' columns as conjugates of these rows)
                f64x4 yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
                if (tw) {
                    const int c0 = (wave % nt) * 16;
                    f64x4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int k0 = kp * (np / KS); k0 < (kp + 1) * (np / KS); k0 += 4) {
                        const int kc = k0 + lq;
                        double tr_ = Xr[lr * ld + kc], ti_ = Xi[lr * ld + kc];
                        if constexpr (KS == 2) {
                            tr_ += Xr[(16 + lr) * ld + kc];
                            ti_ += Xi[(16 + lr) * ld + kc];
                        }
                        const double ur = Ur[kc * ld + c0 + lr], ui = Ui[kc * ld + c0 + lr];
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_, ur, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti_, ui, p2, 0, 0, 0);
                        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr_ + ti_, ur + ui, p3, 0, 0, 0);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        yr[q] = p1[q] - p2[q];
                        yi[q] = p3[q] - (p1[q] + p2[q]);
                    }
                }
                if constexpr (KS == 2) {
                    __syncthreads();
                    if (tw && kp == 1) {
                        const int c0 = (wave % nt) * 16;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            Xr[(lq + 4 * q) * ld + c0 + lr] = yr[q];
                            Xi[(lq + 4 * q) * ld + c0 + lr] = yi[q];
                        }
                    }
                    __syncthreads();
                }
                if (tw && kp == 0) {
                    const int c0 = (wave % nt) * 16;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r = lq + 4 * q, col = c0 + lr;
                        double vr = yr[q], vi = yi[q];
                        if constexpr (KS == 2) {
                            vr += Xr[r * ld + col];
                            vi += Xi[r * ld + col];
                        }
                        if (r < nb_ && col < n) xc[bl[r] * n + col] = make_double2(vr, vi);
                    }
                }
                continue;
            }
            // matrix elements: 16 lanes per (pair, direction), q strided over the lanes, butterfly sum
            const int item = t >> 4, sub = t & 15;
            const int p = item >> 1, d = item & 1;
            const bool ok = p < cnt;
            const int ra = ok ? prow[p][d] : 0, bb = ok ? pband[p][d ^ 1] : 0;
            double sr = 0.0, si = 0.0;
            if (ok) {
                for (int q = sub; q < np; q += 16) {
                    double tr_ = Xr[ra * ld + q], ti_ = Xi[ra * ld + q];
                    if constexpr (KS == 2) {
                        tr_ += Xr[(16 + ra) * ld + q];
                        ti_ += Xi[(16 + ra) * ld + q];
                    }
                    const double ur = Ur[q * ld + bb], ui = Ui[q * ld + bb];
                    sr += tr_ * ur - ti_ * ui;
                    si += tr_ * ui + ti_ * ur;
                }
            }
            for (int o = 8; o > 0; o >>= 1) {
                sr += __shfl_xor(sr, o);
                si += __shfl_xor(si, o);
            }
            if (ok && sub == 0) {
                const int slot = c < WN_CH_A ? c - WN_CH_DDH : 6 + (c - WN_CH_DA);
                pwd[(((size_t)k * WN_PCAP + p) * 2 + d) * 15 + slot] = make_double2(sr, si);
            }
        }
    }
}

// Rotation for n in (64, 128]: U and X are read through the L2 as MFMA operands and U^+ X is staged
// in a per-k scratch [NP][NP] (the LDS cannot hold U, X and U^+ X at these sizes).  One 256-thread
// workgroup per k, channels 1 .. C-1 in turn; the (NP/16)^2 output tiles of each product are dealt to
// the 4 waves.  Complex products as three real products (Gauss): for A B with A = (ar, ai), B = (br, bi),
// P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi); real part P1 - P2, imaginary part P3 - P1 - P2.  Four K
// steps of operands are loaded per iteration, the next iteration's before this one's products.
template <int NP>
__global__ __launch_bounds__(256) void k_rotate_big(double *__restrict__ x, const double2 *__restrict__ u, int n, int C,
                                                   int ldo, double2 *__restrict__ scratch) {
    constexpr int nt = NP / 16;
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const double2 *uk = u + (size_t)k * n * n;
    double2 *tk = scratch + (size_t)k * NP * NP;
    double2 *xk = (double2 *)(x + (size_t)k * ldo);
    for (int c = 1; c < C; ++c) {
        double2 *xc = xk + (size_t)c * n * n;
        // T = U^+ X into the scratch (rows and columns beyond n are zero); (U^+)(r, kk) = conj U(kk, r)
        for (int tile = wave; tile < nt * nt; tile += 4) {
            const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
            f64x4 p1, p2, p3;
            wn_mfma_tile_c64<NP>(uk + min(r0 + lr, n - 1), n, r0 + lr < n, n, -1.0, xc + min(c0 + lr, n - 1), n, c0 + lr < n,
                         n, lq, p1, p2, p3);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                tk[(r0 + lq + 4 * q) * NP + c0 + lr] = make_double2(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
        }
        __threadfence_block();
        __syncthreads();
        // X <- T U
        for (int tile = wave; tile < nt * nt; tile += 4) {
            const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
            f64x4 p1, p2, p3;
            wn_mfma_tile_c64<NP>(tk + (r0 + lr) * NP, 1, true, NP, 1.0, uk + min(c0 + lr, n - 1), n, c0 + lr < n, n, lq, p1, p2,
                         p3);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = r0 + lq + 4 * q, col = c0 + lr;
                if (row < n && col < n) xc[row * n + col] = make_double2(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

size_t wn_eig_scratch_elems(int n, int count) {
    return n <= 32 ? wn_eigd_scratch_elems(n, count) : wn_eig_big_scratch_elems(n, count);
}

size_t wn_rotate_scratch_elems(int n, int count) {
    if (n <= 64) return 0;
    const size_t N = n <= 96 ? 96 : 128;
    return N * N * (size_t)count;
}

template <bool WARM>
static void launch_rotate_mfma(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                               const double2 *uw = nullptr) {
    const int n = sh.n;
    const int np = (n + 15) & ~15;
    const size_t lds = (size_t)4 * np * (np + 1) * sizeof(double);
    const int grid = (WARM && !uw) ? (count + WN_WARM_G - 1) / WN_WARM_G * (WN_WARM_G - 1) : count;
#define WN_ROT(NPV, PFV)                                                                                       \
    hipLaunchKernelGGL((k_rotate_mfma<NPV, PFV, WARM>), dim3(grid), dim3(256), lds, st, x, u, n, sh.C, sh.ldo, count, \
                       uw)
    if (np == 16)
        WN_ROT(16, 1);
    else if (np == 32)
        WN_ROT(32, 4);
    else if (np == 48)
        WN_ROT(48, 9);
    else
        WN_ROT(64, 16);
#undef WN_ROT
}

bool wn_rotate_sc_elements(const WnShape &sh) { return sh.n >= 16 && sh.n <= 64; }

hipError_t wn_launch_rotate_sc(double *x, const double2 *u, const int *count, const int4 *pmeta, double2 *pwd,
                               int nk, const WnShape &sh, hipStream_t st, double2 *scratch, int *flags, bool ch32) {
    if (!wn_rotate_sc_elements(sh) || (ch32 && sh.n % 2 != 0)) {
        if (ch32) return hipErrorInvalidValue;
        return wn_launch_rotate_f64(x, u, nk, sh, st, scratch);
    }
    const int np = (sh.n + 15) & ~15;
    const size_t lds = (size_t)4 * np * (np + 1) * sizeof(double);
#define WN_ROTSC(NPV, PFV)                                                                                          \
    do {                                                                                                            \
        if (ch32)                                                                                                   \
            hipLaunchKernelGGL((k_rotate_sc<NPV, PFV, true>), dim3(nk), dim3(256), lds, st, x, u, count, pmeta, pwd, \
                               sh.n, sh.C, sh.ldo, flags, sh.herm);                                                 \
        else                                                                                                        \
            hipLaunchKernelGGL((k_rotate_sc<NPV, PFV, false>), dim3(nk), dim3(256), lds, st, x, u, count, pmeta,     \
                               pwd, sh.n, sh.C, sh.ldo, flags, sh.herm);                                            \
    } while (0)
    if (np == 16)
        WN_ROTSC(16, 1);
    else if (np == 32)
        WN_ROTSC(32, 4);
    else if (np == 48)
        WN_ROTSC(48, 9);
    else
        WN_ROTSC(64, 16);
#undef WN_ROTSC
    return hipGetLastError();
}

hipError_t wn_launch_eig_f64(const double *x, double *e, double2 *u, int *sweeps, int count, const WnShape &sh,
                             int tpb, hipStream_t st, double2 *vscratch, int *status) {
    return wn_launch_eig_pass_f64(x, e, u, sweeps, nullptr, count, sh, tpb, -1, st, vscratch, nullptr, 0, status);
}

bool wn_warm_start(int n, int count) { return n >= 16 && n <= 128 && count >= WN_WARM_G; }

hipError_t wn_launch_warm_transform_f64(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                                        const double2 *uprev, double2 *vscratch) {
    if (!wn_warm_start(sh.n, count)) return hipErrorInvalidValue;
    if (sh.n > 64) return wn_launch_warm_big(x, u, count, sh, st, uprev, vscratch);
    launch_rotate_mfma<true>(x, u, count, sh, st, uprev);
    return hipGetLastError();
}

hipError_t wn_launch_rotate_f64(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                                double2 *scratch) {
    const int n = sh.n;
    if (n > 128) return hipErrorInvalidValue;
    if (n > 64) {
        if (!scratch) return hipErrorInvalidValue;
        if (n <= 96)
            hipLaunchKernelGGL(k_rotate_big<96>, dim3(count), dim3(256), 0, st, x, u, n, sh.C, sh.ldo, scratch);
        else
            hipLaunchKernelGGL(k_rotate_big<128>, dim3(count), dim3(256), 0, st, x, u, n, sh.C, sh.ldo, scratch);
    } else if (n >= 16) {
        launch_rotate_mfma<false>(x, u, count, sh, st);
    } else {
        size_t lds = (size_t)2 * n * (n + 1) * sizeof(double2);
        hipLaunchKernelGGL(k_rotate_f64, dim3(count), dim3(256), lds, st, x, u, n, sh.C, sh.ldo);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Screening, active pairs and SC/IC kernels per k point (256 threads).
//
// Occupations (CalculationUtils.jl:21-45), df_nm = f_n - f_m, transition energy dE = E_m - E_n.
// Gaussian: pair active iff |df| > 1e-10 and some w_j has |dE - w_j| < sqrt(200) width; the exact
// inclusive window [wlo, whi) of such j is stored.  Lorentzian: pairs with |df| > 1e-10, window
// all j; the k point is active only if some |dE - w_j| < window (TransitionScreening.jl:32-94).
// Pairs are stored n-major.  For each active (n, m) one thread forms the generalized derivatives
// r^b_{nm;a} and r^b_{mn;a} for all 9 (a, b) and writes
//   SC: Q[a+3b+9c] = -i (G_nm[c][a] B_mn[b] - G_mn[b][a] B_nm[c]),   B = A_H + i D
//   IC: Q[27*sc + a+3b+9c] = (V_mm^a - V_nn^a) B_nm[c] B_mn[b].
// ------------------------------------------------------------------------------------------------
// Channels of one k point (channel c at xk + c n^2), addressed by arithmetic rather than a pointer
// table, so that runtime component indices do not force the table into scratch memory.
struct PairView {
    const double2 *xk;
    int n2;
    int has_pos;
    const double2 *pw;   // the pair's dd H and dA elements (k_rotate_sc), per direction; nullptr: read the channels
    unsigned long long bands;   // pair bands of a point whose dH / A rows were stored at them (else 0)
    __device__ __forceinline__ const double2 *V(int y) const { return xk + (size_t)(WN_CH_DH + y) * n2; }
    __device__ __forceinline__ const double2 *W(int s) const { return xk + (size_t)(WN_CH_DDH + s) * n2; }
    __device__ __forceinline__ const double2 *A(int y) const { return xk + (size_t)(WN_CH_A + y) * n2; }
    __device__ __forceinline__ const double2 *dA(int i) const { return xk + (size_t)(WN_CH_DA + i) * n2; }
};

// Lanes per generalized-derivative row (gl = 1, 2, 4 or 8, uniform over the workgroup): the
// intermediate-state sum over l is split into gl strided parts (one per lane of an aligned lane group) and
// added by a butterfly over the group.  Few active pairs get wide groups, so that the items of a chunk
// fill the workgroup in one pass.
#define WN_GL_MAX 8

// Intermediate-state sums of the generalized derivative over this lane's l = sub, sub + gl, ... (all
// bands but n and m):  sAD[y] = sum_l A^x_nl D^y_lm - D^y_nl A^x_lm,  sHD[y] = sum_l V^x_nl D^y_lm - D^y_nl V^x_lm.
// Excluded l get zero weights (D^y_nl = D^y_lm = 0: their products add exact zeros) instead of a branch.
// Column m is read as stored: the interpolated matrices
// need not be Hermitian (the eigensolver reads only the upper triangle of H).
template <bool POS>
__device__ __forceinline__ void wn_gderiv_part(const PairView &pv, const double *inv, int n, int nb, int m, int x,
                                               int sub, int gl, bool ok, double2 sAD[3], double2 sHD[3]) {
    const double2 *Vx = pv.V(x);
    const double2 *Ax = pv.A(x);
    const double2 *V0 = pv.V(0), *V1 = pv.V(1), *V2 = pv.V(2);
    const int lend = ok ? nb : 0;
    for (int l = sub; l < lend; l += gl) {
        const bool skip = (l == n) || (l == m);
        const double inl = skip ? 0.0 : inv[n * nb + l];
        const double ilm = skip ? 0.0 : inv[l * nb + m];
        // column m at row l: for a point whose rotation stored only rows at the pairs' bands (pv.bands != 0) and
        // l not one of them, the element is the conjugate of row m's element l (the channel is Hermitian)
        const bool via_row = pv.bands != 0ull && !((pv.bands >> l) & 1ull);
        const int nl = n * nb + l, lm = via_row ? m * nb + l : l * nb + m;
        const double sg = via_row ? -1.0 : 1.0;
        auto cj = [&](double2 v) { return make_double2(v.x, sg * v.y); };
        const double2 vxnl = Vx[nl], vxlm = cj(Vx[lm]);
        const double2 vnl[3] = {V0[nl], V1[nl], V2[nl]};
        const double2 vlm[3] = {cj(V0[lm]), cj(V1[lm]), cj(V2[lm])};
        double2 axnl = make_double2(0.0, 0.0), axlm = make_double2(0.0, 0.0);
        if (POS) {
            axnl = Ax[nl];
            axlm = cj(Ax[lm]);
        }
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            const double2 dnl = cscale(vnl[y], -inl);
            const double2 dlm = cscale(vlm[y], -ilm);
            if (POS) sAD[y] = csub(cfma(axnl, dlm, sAD[y]), cmul(dnl, axlm));
            sHD[y] = csub(cfma(vxnl, dlm, sHD[y]), cmul(dnl, vxlm));
        }
    }
}

// Row r^y_{nm;x}, y = 0..2, of the generalized derivative, computed by the gl lanes of an aligned lane
// group (sub = lane index in the group; every lane of the group must call it; lanes with ok = false add
// nothing).  Every lane of the group returns the row; the caller stores it from one lane.
__device__ void wn_gderiv_row(const PairView &pv, const double *inv, int n, int nb, int m, int x, double2 g[3],
                              int sub, int gl, bool ok) {
    double2 sAD[3], sHD[3];
#pragma unroll
    for (int y = 0; y < 3; ++y) {
        sAD[y] = make_double2(0.0, 0.0);
        sHD[y] = make_double2(0.0, 0.0);
    }
    if (pv.has_pos)
        wn_gderiv_part<true>(pv, inv, n, nb, m, x, sub, gl, ok, sAD, sHD);
    else
        wn_gderiv_part<false>(pv, inv, n, nb, m, x, sub, gl, ok, sAD, sHD);
    for (int o = gl / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            sAD[y].x += __shfl_xor(sAD[y].x, o);
            sAD[y].y += __shfl_xor(sAD[y].y, o);
            sHD[y].x += __shfl_xor(sHD[y].x, o);
            sHD[y].y += __shfl_xor(sHD[y].y, o);
        }
    }
    const double2 *Vx = pv.V(x);
    const double2 *Ax = pv.A(x);
    const double inm = inv[n * nb + m];
    double2 Dnm[3], dA[3], dV[3];
#pragma unroll
    for (int y = 0; y < 3; ++y) {
        Dnm[y] = cscale(pv.V(y)[n * nb + m], -inm);
        dV[y] = csub(pv.V(y)[n * nb + n], pv.V(y)[m * nb + m]);
        dA[y] = pv.has_pos ? csub(pv.A(y)[n * nb + n], pv.A(y)[m * nb + m]) : make_double2(0.0, 0.0);
    }
    const double2 Anm = pv.has_pos ? pv.A(x)[n * nb + m] : make_double2(0.0, 0.0);
    // the x entries, read again rather than indexed with the runtime x (which would put the arrays in
    // scratch memory)
    const double2 Dnm_x = cscale(Vx[n * nb + m], -inm);
    const double2 dV_x = csub(Vx[n * nb + n], Vx[m * nb + m]);
    const double2 dA_x = pv.has_pos ? csub(Ax[n * nb + n], Ax[m * nb + m]) : make_double2(0.0, 0.0);
#pragma unroll
    for (int y = 0; y < 3; ++y) {
        const double2 ad_bit = cadd(cmul(dA[y], Dnm_x), cmul(dA_x, Dnm[y]));
        const double2 aa_bit = cmul(dA[y], Anm);
        const double2 dv_bit = cadd(cmul(dV[y], Dnm_x), cmul(dV_x, Dnm[y]));
        const double2 da = pv.has_pos ? (pv.pw ? pv.pw[6 + 3 * x + y] : pv.dA(3 * x + y)[n * nb + m])
                                      : make_double2(0.0, 0.0);
        const double2 w = pv.pw ? pv.pw[wn_sym(x, y)] : pv.W(wn_sym(x, y))[n * nb + m];
        double2 gg = cadd(da, ad_bit);
        gg = csub(gg, cmuli(aa_bit));
        gg = cadd(gg, sAD[y]);
        const double2 h = cadd(cadd(w, sHD[y]), dv_bit);
        g[y] = cadd(gg, cscale(cmuli(h), -inm));
    }
}

// Photon-energy window of a transition energy de: the first and last j with |de - w_j| < cutoff
// (Gaussian), found from the first w_j above de - cutoff; first = -1 when there is none.  A transition
// with (w_0 - de) >= cutoff or (de - w_last) >= cutoff has none (rounded differences are monotone in w_j).
__device__ __forceinline__ void wn_gauss_window(double de, const double *__restrict__ omega, int nw, double cutoff,
                                                double om_first, double om_last, int &first, int &last) {
    first = -1;
    last = -1;
    if (om_first - de >= cutoff || de - om_last >= cutoff) return;
    int lo = 0, hi = nw;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (omega[mid] > de - cutoff)
            hi = mid;
        else
            lo = mid + 1;
    }
    for (int j = lo > 0 ? lo - 1 : 0; j < nw; ++j) {
        const double xx = de - omega[j];
        if (fabs(xx) < cutoff) {
            if (first < 0) first = j;
            last = j;
        } else if (omega[j] > de) {
            break;
        }
    }
}

// Screening and active-pair lists, one wave per k point (four k points per workgroup, no block
// barriers).  Occupations (CalculationUtils.jl:21-45), df_nm = f_n - f_m, dE = E_m - E_n; candidates
// (n, m) in n-major order, 64 per round, kept pairs compacted in order with a ballot:
//   Gaussian: |df| > 1e-10 and some w_j with |dE - w_j| < sqrt(200) width; window [first, last + 1).
//   Lorentzian: |df| > 1e-10, window all j; the k point is active only if some |dE - w_j| < window.
// (TransitionScreening.jl:32-94.)  count[k] = number of pairs (0 for an inactive point); with
// denominator_regularization = 0 an active point with two levels closer than 1e-10 sets flags[0].
__global__ __launch_bounds__(256) void k_screen(const double *__restrict__ eg, const double *__restrict__ omega,
                                               int *__restrict__ count, int4 *__restrict__ pmeta,
                                               double2 *__restrict__ pmd, int *__restrict__ flags, int nk, WnShape sh,
                                               WnScreen sc) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nk) return;
    const int nb = sh.n;
    const int n2 = nb * nb;
    const double el = lane < nb ? eg[(size_t)k * nb + lane] : 0.0;
    double fl = 0.0;
    {
        const double kt = 8.617333262e-5 * sc.temperature;
        if (kt == 0.0) {
            fl = el <= sc.mu ? 1.0 : 0.0;
        } else {
            const double arg = fmin(fmax((el - sc.mu) / kt, -700.0), 700.0);
            fl = 1.0 / (exp(arg) + 1.0);
        }
    }
    const double cutoff = sc.width * sqrt(200.0);
    const double om_first = sc.nw > 0 ? omega[0] : 0.0, om_last = sc.nw > 0 ? omega[sc.nw - 1] : 0.0;
    const size_t pbase = (size_t)k * n2;
    int total = 0;
    bool active = false;
    for (int c0 = 0; c0 < n2; c0 += 64) {
        const int c = c0 + lane;
        const int n = c < n2 ? c / nb : 0, m = c < n2 ? c - n * nb : 0;
        const double en = __shfl(el, n), em = __shfl(el, m);
        const double df = __shfl(fl, n) - __shfl(fl, m);
        const double de = em - en;
        bool keep = c < n2 && fabs(df) > 1.0e-10;
        bool flag = false;
        int first = 0, last = sc.nw - 1;
        if (keep) {
            if (sc.gaussian) {
                wn_gauss_window(de, omega, sc.nw, cutoff, om_first, om_last, first, last);
                keep = first >= 0;
                flag = keep;
            } else if (sc.window > 0.0) {
                for (int j = 0; j < sc.nw; ++j)
                    if (fabs(de - omega[j]) < sc.window) {
                        flag = true;
                        break;
                    }
            } else {
                flag = true;
            }
        }
        active = active || __ballot(flag) != 0ull;
        const unsigned long long kept = __ballot(keep);
        if (keep) {
            const int slot = total + __popcll(kept & ((1ull << lane) - 1ull));
            pmeta[pbase + slot] = make_int4(n, m, first, last + 1);
            pmd[pbase + slot] = make_double2(de, df);
        }
        total += __popcll(kept);
    }
    if (!active) total = 0;
    if (lane == 0) count[k] = total;
    // running count of active pairs (counter buckets) for the work figures of the response stage
    if (lane == 0 && total > 0)
        atomicAdd((unsigned long long *)(flags + WN_CTR_BASE + 4 + WN_CTR_STRIDE * (k % WN_CTR_BUCKETS)),
                  (unsigned long long)total);
    // eta == 0: bare inverse requires isolated levels (OrthogonalBandDerivatives.jl:66-82)
    if (active && sc.eta == 0.0) {
        bool bad = false;
        for (int j = 0; j < nb; ++j) {
            const double ej = __shfl(el, j);
            if (lane < nb && j > lane && !(fabs(ej - el) > 1.0e-10)) bad = true;
        }
        if (bad) atomicOr(flags, 1);
    }
}

// Generalized derivatives and SC/IC pair kernels of the pairs k_screen listed, one 256-thread
// workgroup per k point.
__global__ __launch_bounds__(256, 3) void k_pairs_f64(const double *__restrict__ x, const double *__restrict__ eg,
                                                  const int *__restrict__ count, const int4 *__restrict__ pmeta,
                                                  const double2 *__restrict__ pwd, double2 *__restrict__ qout,
                                                  WnShape sh, WnScreen sc) {
    __shared__ double E[64];
    extern __shared__ double inv[];   // [n][n] regularized inverse energy differences
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    const int nb = sh.n;
    const int n2 = nb * nb;
    const int total = count[k];
    if (total == 0) return;
    if (t < nb) E[t] = eg[(size_t)k * nb + t];
    __syncthreads();
    const size_t pbase = (size_t)k * n2;
    for (int e = t; e < n2; e += 256) {
        const int i = e / nb, j = e - i * nb;
        inv[e] = wn_reg_inverse(E[i] - E[j], sc.eta);
    }
    __syncthreads();

    const double2 *xk = (const double2 *)(x + (size_t)k * sh.ldo);
    PairView pv;
    pv.xk = xk;
    pv.n2 = n2;
    pv.has_pos = sh.has_pos;
    pv.pw = nullptr;
    pv.bands = 0ull;
    // dd H and dA elements precomputed by k_rotate_sc for points with at most WN_PCAP pairs
    const bool compact = pwd != nullptr && total <= WN_PCAP;
    // Hermitian model: k_rotate_sc stored the velocity and position channels of these points as rows at the
    // pairs' bands only (n <= 64); element (l, m) at another band l is the conjugate of row m's element l
    __shared__ unsigned long long bmk;
    if (compact && sh.herm) {
        if (t == 0) bmk = 0ull;
        __syncthreads();
        if (t < total) {
            const int4 mt = pmeta[pbase + t];
            atomicOr(&bmk, (1ull << mt.x) | (1ull << mt.y));
        }
        __syncthreads();
        pv.bands = bmk;
    }
    const int ncomp = 27 * sc.nq;
    // Chunks of WN_PCH pairs.  Phase A: one item per (pair, direction d, component x) forms
    // r^y_{n'm';x} for y = 0..2 with (n', m') = (n, m) or (m, n); one item per pair also stages the
    // Berry connections B_nm, B_mn and the velocity differences.  Phase B: one item per (pair, output).
    __shared__ double2 Gs[WN_PCH][2][3][3];
    __shared__ double2 Bs[WN_PCH][2][3];
    __shared__ double2 Vd[WN_PCH][3];
    for (int c0 = 0; c0 < total; c0 += WN_PCH) {
        const int cn = min(WN_PCH, total - c0);
        // gl lanes per item; the loop bound is uniform so that every lane group reaches the butterfly
        const int gl = cn * 6 <= 256 / WN_GL_MAX ? WN_GL_MAX : (cn * 6 <= 64 ? 4 : (cn * 6 <= 128 ? 2 : 1));
        for (int ib = 0; ib < cn * 6; ib += 256 / gl) {
            const int item = ib + t / gl, sub = t % gl;
            const bool ok = item < cn * 6;
            const int i = ok ? item % cn : 0, dx = ok ? item / cn : 0;
            const int d = dx / 3, xcomp = dx - 3 * d;
            const int4 mt = pmeta[pbase + c0 + i];
            const int n = d ? mt.y : mt.x, m = d ? mt.x : mt.y;
            double2 g[3];
            PairView pvi = pv;
            if (compact) pvi.pw = pwd + (((size_t)k * WN_PCAP + c0 + i) * 2 + d) * 15;
            wn_gderiv_row(pvi, inv, n, nb, m, xcomp, g, sub, gl, ok);
            if (!ok || sub != 0) continue;
#pragma unroll
            for (int y = 0; y < 3; ++y) Gs[i][d][xcomp][y] = g[y];
            if (dx == 0) {
                const double inm = inv[n * nb + m];
                const double imn = inv[m * nb + n];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double2 anm = sh.has_pos ? pv.A(a)[n * nb + m] : make_double2(0.0, 0.0);
                    const double2 amn = sh.has_pos ? pv.A(a)[m * nb + n] : make_double2(0.0, 0.0);
                    const double2 dnm = cscale(pv.V(a)[n * nb + m], -inm);
                    const double2 dmn = cscale(pv.V(a)[m * nb + n], -imn);
                    Bs[i][0][a] = cadd(anm, cmuli(dnm));
                    Bs[i][1][a] = cadd(amn, cmuli(dmn));
                    Vd[i][a] = csub(pv.V(a)[m * nb + m], pv.V(a)[n * nb + n]);
                }
            }
        }
        __syncthreads();
        for (int item = t; item < cn * ncomp; item += 256) {
            const int i = item / ncomp, comp = item - i * ncomp;
            double2 *q = qout + (pbase + c0 + i) * ncomp;
            const int sc_on = sc.sc;
            if (sc_on && comp < 27) {
                const int a = comp % 3, b = (comp / 3) % 3, c = comp / 9;
                // -i (G_nm[c][a] B_mn[b] - G_mn[b][a] B_nm[c])
                const double2 z = csub(cmul(Gs[i][0][c][a], Bs[i][1][b]), cmul(Gs[i][1][b][a], Bs[i][0][c]));
                q[comp] = make_double2(z.y, -z.x);
            } else {
                const int cc = comp - (sc_on ? 27 : 0);
                const int a = cc % 3, b = (cc / 3) % 3, c = cc / 9;
                q[comp] = cmul(cmul(Vd[i][a], Bs[i][0][c]), Bs[i][1][b]);
            }
        }
        __syncthreads();
    }
}

hipError_t wn_launch_screen(const double *e, const double *omega, int *count, int4 *pmeta, double2 *pmd, int *flags,
                            int nk, const WnShape &sh, const WnScreen &sc, hipStream_t st) {
    if (sh.n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_screen, dim3((nk + 3) / 4), dim3(256), 0, st, e, omega, count, pmeta, pmd, flags, nk, sh, sc);
    return hipGetLastError();
}

hipError_t wn_launch_pairs_f64(const double *x, const double *e, const int *count, const int4 *pmeta,
                               const double2 *pwd, double2 *q, int nk, const WnShape &sh, const WnScreen &sc,
                               hipStream_t st) {
    if (sh.n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pairs_f64, dim3(nk), dim3(256), (size_t)sh.n * sh.n * sizeof(double), st, x, e, count, pmeta,
                       pwd, q, sh, sc);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Point weights folded into the pairs' occupation differences (one workgroup of 64 per point).
__global__ __launch_bounds__(64) void k_weight_pairs(double2 *__restrict__ pmd, const int *__restrict__ count,
                                                     const double *__restrict__ w, int pcap) {
    const int k = blockIdx.x;
    const int c = count[k];
    const double wk = w[k];
    double2 *row = pmd + (size_t)k * pcap;
    for (int j = threadIdx.x; j < c; j += 64) row[j].y *= wk;
}

hipError_t wn_launch_weight_pairs(double2 *pmd, const int *count, const double *w, int nk, const WnShape &sh,
                                  hipStream_t st) {
    if (nk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_weight_pairs, dim3(nk), dim3(64), 0, st, pmd, count, w, sh.n * sh.n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Frequency contraction per chunk of k points.  Workgroup (chunk, 64-frequency tile): wave w
// owns components [14w, 14w+14), lane = frequency.  The chunk's pairs are staged in LDS in blocks of
// 256 in (k, pair) order, those whose frequency window meets the tile are compacted (order kept), and
// for each of them weight = df * delta(dE - w) (exact Gaussian cutoff test per frequency) and
// acc[c] += weight * Q[c]; the per-chunk sums go to
// partial[chunk][component][frequency].  The summation order is the (k, pair) order.
__global__ __launch_bounds__(256) void k_accum(const int4 *__restrict__ chunks, const int *__restrict__ count,
                                              const int4 *__restrict__ pmeta, const double2 *__restrict__ pmd,
                                              const double2 *__restrict__ q, const double *__restrict__ omega,
                                              double2 *__restrict__ partial, int pcap, int ncomp, int nw, int nwt,
                                              int gaussian, double width) {
    __shared__ int scnt[WN_ACC_KMAX + 1];
    __shared__ int wsel[4];
    __shared__ double2 smd[256];
    __shared__ long long srow[256];
    const int chunk = blockIdx.x / nwt;
    const int wt = blockIdx.x - chunk * nwt;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int w = wt * WN_WT + lane;
    const double om = w < nw ? omega[w] : 0.0;
    const int cbase = wave * 14;
    const double cutoff = width * sqrt(200.0);
    const double gscale = 1.0 / (sqrt(M_PI) * width);
    const double lscale = width / M_PI;
    const int tlo = wt * WN_WT, thi = tlo + WN_WT;
    double2 acc[14];
#pragma unroll
    for (int c = 0; c < 14; ++c) acc[c] = make_double2(0.0, 0.0);
    const int4 ch = chunks[chunk];
    if (t == 0) {
        int run = 0;
        for (int j = 0; j < ch.y; ++j) {
            scnt[j] = run;
            run += count[ch.x + j];
        }
        scnt[ch.y] = run;
    }
    __syncthreads();
    const int total = scnt[ch.y];
    for (int b0 = 0; b0 < total; b0 += 256) {
        const int nbk = min(256, total - b0);
        // stage pair b0 + t, keep it if its window [z, w) meets the tile
        bool keep = false;
        long long row = 0;
        double2 md = make_double2(0.0, 0.0);
        if (t < nbk) {
            const int j = b0 + t;
            int kk = 0;
            while (scnt[kk + 1] <= j) ++kk;
            row = (long long)(ch.x + kk) * pcap + (j - scnt[kk]);
            const int4 mt = pmeta[row];
            keep = !(mt.w <= tlo || mt.z >= thi);
            if (keep) md = pmd[row];
        }
        // ordered compaction of the kept pairs
        const unsigned long long bal = __ballot(keep);
        if (lane == 0) wsel[wave] = __popcll(bal);
        __syncthreads();
        int off = 0;
        for (int v = 0; v < wave; ++v) off += wsel[v];
        const int nsel = wsel[0] + wsel[1] + wsel[2] + wsel[3];
        if (keep) {
            const int pos = off + __popcll(bal & ((1ull << lane) - 1ull));
            smd[pos] = md;
            srow[pos] = row;
        }
        __syncthreads();
        for (int sidx = 0; sidx < nsel; ++sidx) {
            const double2 pm = smd[sidx];
            const double2 *qq = q + srow[sidx] * ncomp + cbase;
            double2 v[14];
#pragma unroll
            for (int c = 0; c < 14; ++c) v[c] = cbase + c < ncomp ? qq[c] : make_double2(0.0, 0.0);
            const double xx = pm.x - om;
            double dl;
            if (gaussian) {
                dl = fabs(xx) < cutoff ? gscale * exp(-(xx / width) * (xx / width)) : 0.0;
            } else {
                dl = lscale / (xx * xx + width * width);
            }
            const double wgt = w < nw ? pm.y * dl : 0.0;
#pragma unroll
            for (int c = 0; c < 14; ++c) {
                acc[c].x += wgt * v[c].x;
                acc[c].y += wgt * v[c].y;
            }
        }
        __syncthreads();
    }
    if (w < nw) {
        double2 *o = partial + (size_t)chunk * ncomp * nw;
#pragma unroll
        for (int c = 0; c < 14; ++c)
            if (cbase + c < ncomp) o[(size_t)(cbase + c) * nw + w] = acc[c];
    }
}

hipError_t wn_launch_accum(const int4 *chunks, int nchunks, const int *count, const int4 *pmeta, const double2 *pmd,
                           const double2 *q, const double *omega, double2 *partial, const WnShape &sh,
                           const WnScreen &sc, hipStream_t st) {
    const int nwt = (sc.nw + WN_WT - 1) / WN_WT;
    const int ncomp = 27 * sc.nq;
    if (ncomp > 56) return hipErrorInvalidValue;   // four waves of 14 components
    hipLaunchKernelGGL(k_accum, dim3(nchunks * nwt), dim3(256), 0, st, chunks, count, pmeta, pmd, q, omega, partial,
                       sh.n * sh.n, ncomp, sc.nw, nwt, sc.gaussian, sc.width);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Chunk partials -> lanes in fixed chunk order (Kahan-compensated when lanecomp != nullptr).
// lanes[i] = (lane, chunk_begin, chunk_end, -).
// ------------------------------------------------------------------------------------------------
// Chunk partials -> lane accumulators in a fixed order.  A workgroup owns WN_LR_E consecutive
// elements of one lane; its WN_LR_S segments of threads sum the lane's chunks c = first + s,
// first + s + WN_LR_S, ... in increasing c, and segment sums are added to the lane in segment order,
// so the result depends only on the chunk plan.  With lanecomp (the mixed forms) every addition is
// Kahan-compensated.
#define WN_LR_E 16
#define WN_LR_S 16
__global__ __launch_bounds__(WN_LR_E * WN_LR_S) void k_lane_reduce(const int4 *__restrict__ lanes,
                                                                  const double2 *__restrict__ partial,
                                                                  double2 *__restrict__ lanebuf,
                                                                  double2 *__restrict__ lanecomp, int nelem) {
    __shared__ double2 ssum[WN_LR_S][WN_LR_E];
    __shared__ double2 scomp[WN_LR_S][WN_LR_E];
    const int j = threadIdx.x % WN_LR_E, sg = threadIdx.x / WN_LR_E;
    const int e = blockIdx.x * WN_LR_E + j;
    const int4 li = lanes[blockIdx.y];
    double2 acc = make_double2(0.0, 0.0), comp = make_double2(0.0, 0.0);
    if (e < nelem) {
        for (int c = li.y + sg; c < li.z; c += WN_LR_S) {
            const double2 v = partial[(size_t)c * nelem + e];
            if (lanecomp) {
                const double yx = v.x - comp.x, yy = v.y - comp.y;
                const double tx = acc.x + yx, ty = acc.y + yy;
                comp.x = (tx - acc.x) - yx;
                comp.y = (ty - acc.y) - yy;
                acc.x = tx;
                acc.y = ty;
            } else {
                acc.x += v.x;
                acc.y += v.y;
            }
        }
    }
    ssum[sg][j] = acc;
    scomp[sg][j] = comp;
    __syncthreads();
    if (sg != 0 || e >= nelem) return;
    const size_t o = (size_t)li.x * nelem + e;
    double2 lane = lanebuf[o];
    if (lanecomp) {
        double2 lc = lanecomp[o];
        for (int q = 0; q < WN_LR_S; ++q) {
            // segment value sum - compensation, added with compensation
            const double2 sv = ssum[q][j], sc = scomp[q][j];
            const double vx = sv.x - sc.x, vy = sv.y - sc.y;
            const double yx = vx - lc.x, yy = vy - lc.y;
            const double tx = lane.x + yx, ty = lane.y + yy;
            lc.x = (tx - lane.x) - yx;
            lc.y = (ty - lane.y) - yy;
            lane.x = tx;
            lane.y = ty;
        }
        lanecomp[o] = lc;
    } else {
        for (int q = 0; q < WN_LR_S; ++q) {
            lane.x += ssum[q][j].x;
            lane.y += ssum[q][j].y;
        }
    }
    lanebuf[o] = lane;
}

hipError_t wn_launch_lane_reduce(const int4 *lanes, int nlanes, const double2 *partial, double2 *lanebuf,
                                 double2 *lanecomp, int nelem, hipStream_t st) {
    if (nlanes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lane_reduce, dim3((nelem + WN_LR_E - 1) / WN_LR_E, nlanes), dim3(WN_LR_E * WN_LR_S), 0, st,
                       lanes, partial, lanebuf, lanecomp, nelem);
    return hipGetLastError();
}
