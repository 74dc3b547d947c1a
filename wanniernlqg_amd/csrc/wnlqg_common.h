// Shared device helpers of the MI355X k-mesh response engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// MFMA lane layouts used here (measured on gfx950):
//   v_mfma_f64_16x16x4f64   A: lane l -> A[l%16][l/16];  B: B[l/16][l%16];  D reg q -> D[l/16 + 4q][l%16]
//   v_mfma_f32_32x32x16_bf16 A: A[l%32][8(l/32)+j]; B: B[8(l/32)+j][l%32];
//                            D reg q -> D[8(q/4) + 4(l/32) + q%4][l%32]
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Interpolation channel map.  Channel c of the per-k block [C][n][n]:
//   0 H, 1-3 dH_a, 4-9 ddH for (xx,yy,zz,yz,xz,xy), 10-12 A_a, 13-21 dA_ab (13 + 3a + b).
#define WN_CH_H 0
#define WN_CH_DH 1
#define WN_CH_DDH 4
#define WN_CH_A 10
#define WN_CH_DA 13
#define WN_CH_BASE 10
#define WN_CH_FULL 22

// GEMM tile of the interpolation kernels: 128 k points x 128 real output columns,
// K blocks of 8 R vectors (16 real rows: 8 cos parts then 8 sin parts).
#define WN_GM 128
#define WN_GN 128
#define WN_RB 8

// k points of one accumulation chunk and frequencies of one accumulation tile.
#define WN_WT 64

__host__ __device__ inline int wn_sym(int a, int b) {
    // (0,0)->0 (1,1)->1 (2,2)->2 (1,2)->3 (0,2)->4 (0,1)->5
    if (a == b) return a;
    int s = a + b;  // 3 -> (1,2), 2 -> (0,2), 1 -> (0,1)
    return s == 3 ? 3 : (s == 2 ? 4 : 5);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// i * a
__device__ __forceinline__ double2 cmuli(double2 a) { return make_double2(-a.y, a.x); }
// a*b + c
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
    return make_double2(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}

// d/(d^2+eta^2) with the scaled branch for over/underflow
// (reference PositionVertexRegularization.jl:17-29); eta == 0 selects the bare 1/d (0 at d == 0).
// 1/sqrt(x) and 1/x for finite x > 0: hardware estimate, then two Newton steps (quadratic
// convergence takes the estimate's error far below the FP64 rounding level).
__device__ __forceinline__ double wn_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}
__device__ __forceinline__ double wn_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = y * (2.0 - x * y);
    y = y * (2.0 - x * y);
    return y;
}

__device__ __forceinline__ double wn_reg_inverse(double d, double eta) {
    if (eta == 0.0) return d == 0.0 ? 0.0 : 1.0 / d;
    double den = d * d + eta * eta;
    if (isfinite(den) && den != 0.0) return d / den;
    double s = fmax(fabs(d), eta);
    double sd = d / s, sw = eta / s;
    return (sd / (sd * sd + sw * sw)) / s;
}

// One 16 x 16 tile of a complex product A B over K = NP on FP64 MFMA (three real products per complex product).  Lane (lr, lq) reads A at a[kk * a_ks] (this lane's
// row; valid while a_ok and kk < a_kmax; imaginary part times a_im) and B at b[kk * b_ks] (this lane's
// column; valid while b_ok and kk < b_kmax), kk = k0 + lq.
template <int NP>
__device__ __forceinline__ void wn_mfma_tile_c64(const double2 *__restrict__ a, int a_ks, bool a_ok, int a_kmax, double a_im,
                                         const double2 *__restrict__ b, int b_ks, bool b_ok, int b_kmax, int lq,
                                         f64x4 &p1, f64x4 &p2, f64x4 &p3) {
    const double2 zero = make_double2(0.0, 0.0);
    p1 = (f64x4){0.0, 0.0, 0.0, 0.0};
    p2 = p1;
    p3 = p1;
    double2 av[4], bv[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
        const int kk = 4 * s4 + lq;
        av[s4] = (a_ok && kk < a_kmax) ? a[kk * a_ks] : zero;
        bv[s4] = (b_ok && kk < b_kmax) ? b[kk * b_ks] : zero;
    }
#pragma unroll
    for (int k0 = 0; k0 < NP; k0 += 16) {
        double2 an[4], bn[4];
        if (k0 + 16 < NP) {
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int kk = k0 + 16 + 4 * s4 + lq;
                an[s4] = (a_ok && kk < a_kmax) ? a[kk * a_ks] : zero;
                bn[s4] = (b_ok && kk < b_kmax) ? b[kk * b_ks] : zero;
            }
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const double ar = av[s4].x, ai = a_im * av[s4].y;
            const double2 bq = bv[s4];
            p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bq.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bq.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, bq.x + bq.y, p3, 0, 0, 0);
        }
        if (k0 + 16 < NP) {
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                av[s4] = an[s4];
                bv[s4] = bn[s4];
            }
        }
    }
}
