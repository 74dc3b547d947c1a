// Launch wrappers of the engine's HIP kernels (host-callable, C++ linkage, internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wnlqg_common.h"

struct WnShape {
    int n;        // orbitals
    int nr;       // R vectors
    int C;        // interpolation channels stored per k point
    int has_pos;  // position channels present
    int Kb;       // blocks of 8 R vectors
    int nNt;      // 128-column tiles of the interpolation output
    int ldo;      // per-k row stride (doubles) of the interpolation output = nNt * 128
    int S;        // span of one phase table axis
    int max_sweeps;  // Jacobi sweep cap; a solve that reaches it sets status word 1
    int rec_sweeps;  // deferred-eigenvector Jacobi (n <= 32): sweeps recorded before the continuation kernel
    int herm;        // the model is exactly Hermitian (H(-R) = H(R)^+, A(-R) = A(R)^+, equal degeneracies), so
                     // every interpolated channel is Hermitian and the SC/IC element rotation may form rows only
    // stored channel s holds logical channel lch[s] of the full map (WN_CH_H .. WN_CH_DA + 8)
    signed char lch[WN_CH_FULL];
};

// Berry-curvature dipole / intrinsic anomalous Hall per-k stage.
struct WnGeom {
    int kind;            // 1 BCD (Fermi-surface moments), 2 AHC (Matrix method, T = 0)
    int nmu;             // Fermi energies (<= WN_GEOM_MAXMU)
    double temperature;  // K
    double eta_fs;       // Gaussian width of -f'(E) at T = 0 (BCD)
    double eta;          // denominator regularization of D (0: bare 1/dE)
    double gap_tol;      // BCD spectral isolation threshold (eV)
    int c_dh;            // stored slot of dH_x (dH_y, dH_z follow)
    int c_a;             // stored slot of A_x, or -1 without a position operator
    int c_da;            // stored slot of dA_xx (d_b A_a at c_da + 3a + b)
};
#define WN_GEOM_MAXMU 64

// Second-harmonic generation parameters (SecondHarmonicKernels.jl:115-236, Config.jl:107-109).
struct WnShg {
    int ncomp;             // requested components (1 .. 18)
    int comp[18];          // a + 3b + 9c of each component
    int nw;                // photon energies
    int gaussian;          // 1 Gaussian, 0 Lorentzian
    double width;          // broadening (eV)
    double lowb;           // low-frequency broadening (eV)
    double eta;            // intermediate regularization (eV)
    int eta_correction;    // 1: eta-corrected optical position derivative
    double mu;             // Fermi energy
    double temperature;    // K
    double tol;            // degeneracy threshold of the one-band BCD term
};

struct WnScreen {
    int nw;             // photon energies
    int gaussian;       // 1 Gaussian, 0 Lorentzian
    double width;       // broadening
    double mu;          // Fermi energy
    double temperature; // K
    double eta;         // denominator regularization
    double window;      // transition_window_factor * width (Lorentzian screening)
    int nq;             // states (1 or 2)
    int sc;             // shift current requested
    int ic;             // injection current requested
};

// Build the interpolation B operand in fragment order from H(R), A(R) (Julia layouts).
hipError_t wn_launch_build_coeff(const double *hr, const double *ar, const double *rc, const int *deg,
                                 double *bf, const WnShape &sh, hipStream_t st);
// k points of a mesh batch: global indices k0 .. k0+count-1, rows up to rows_pad (extra rows k = 0).
hipError_t wn_launch_mesh_kpoints(double *kp, const int *mesh, long long k0, int count, int rows_pad,
                                  hipStream_t st);
// Phase tables T[k][axis][j] = exp(2 pi i k_axis (rmin_axis + j)) (FP64, and FP32 if tab32 != nullptr).
hipError_t wn_launch_phase_tables(const double *kp, const int *rmin, double2 *tab, float2 *tab32, int rows, int S,
                                  hipStream_t st);
// Split-BF16 coefficient operand (hi, mid and optional lo parts) for the mixed interpolation.
hipError_t wn_launch_build_coeff_bf16(const double *hr, const double *ar, const double *rc, const int *deg,
                                      __bf16 *bh, __bf16 *bm, __bf16 *bl, const WnShape &sh, hipStream_t st);
// FP64 MFMA interpolation GEMM over the sh.nNt 128-column tiles stored in bf.
hipError_t wn_launch_interp_f64(const double *bf, const double2 *tab, const int4 *ridx, double *out, int rows_pad,
                                const WnShape &sh, hipStream_t st);
// Split-BF16 MFMA interpolation GEMM (FP32 accumulate, FP64 output); nterms = 3 or 6.  Computes the
// 256-column tiles from the one containing real column col0 to the end of the row.
hipError_t wn_launch_interp_bf16(const __bf16 *bh, const __bf16 *bm, const __bf16 *bl, const float2 *tab,
                                 const int4 *ridx, double *out, int rows_pad, int nterms, int col0, const WnShape &sh,
                                 hipStream_t st);
// Split-FP16 mixed interpolation: per-column scales, FP16 coefficient parts, (R1,R2)-plane phase tables and GEMM.
hipError_t wn_launch_build_coeff_f16(const double *hr, const double *ar, const double *rc, const int *deg,
                                     double *up, double *down, _Float16 *bh, _Float16 *bm, const WnShape &sh,
                                     hipStream_t st);
hipError_t wn_launch_phase_planes(const double2 *tab, float2 *t12, float2 *t3, int rows, int S, hipStream_t st);
hipError_t wn_launch_interp_f16(const _Float16 *bh, const _Float16 *bm, const double *cscale, const float2 *t12,
                                const float2 *t3, const int2 *ridx2, double *out, int rows_pad, int nterms, int col0,
                                const WnShape &sh, hipStream_t st);
// Separable mesh interpolation: dense R box of channel rows, per-axis phases, small complex contraction.
hipError_t wn_launch_build_box(const double *hr, const double *ar, const double *rc, const int *deg, const int *boxmap,
                               double *box, long long nbox, const WnShape &sh, hipStream_t st);
hipError_t wn_launch_axis_phases(double2 *E, int N, int span, int rmin, hipStream_t st);
// n2c32 > 0 (last stage of the mixed form): entries m >= n2c32 of an output row stored as FP32, channel
// c = m / n2c32 in the first half of its FP64 slot.
hipError_t wn_launch_contract(const double2 *in, double2 *out, const double2 *E, int P, int p0, int Q, int M, int nb,
                              long long ib, long long iq, long long ob, long long op, hipStream_t st, int n2c32 = 0);
// Selected outputs of a contraction: group g = (b, t0, t1, -) writes out[t*ot + m] for t0 <= t < t1 from input
// row b and phase row sel[t] (one mesh row's chosen points, written compactly).
hipError_t wn_launch_contract_sel(const double2 *in, double2 *out, const double2 *E, const int4 *groups, int ngroups,
                                  const int *sel, int Q, int M, long long ib, long long iq, long long ot,
                                  hipStream_t st);
// Jacobi eigensolver of the upper triangle of channel 0.
// tpb: threads per k point (64, 128, 256; 0 = default of the size class).
// vscratch: wn_eig_scratch_elems(n, count) complex elements, required for n > 64 (deferred-eigenvector solver).
// status (nullable): word 1 |= 1 when a point reaches sh.max_sweeps without meeting the stopping test.
hipError_t wn_launch_eig_f64(const double *x, double *e, double2 *u, int *sweeps, int count, const WnShape &sh,
                             int tpb, hipStream_t st, double2 *vscratch = nullptr, int *status = nullptr);
// Warm-started solve of count consecutive points of a mesh row (groups of 4 along the fastest axis):
//   pass 0  wn_launch_eig_pass_f64(..., pass = 0): the group anchors from the identity;
//           wn_launch_warm_transform_f64: channel 0 of the other points <- U_a^+ H U_a;
//   pass 1  wn_launch_eig_pass_f64(..., pass = 1): the other points, V started at U_a.
// pass = -1 solves every point from the identity.  wn_warm_start(n, count) tells whether the warm
// sequence applies (n >= 16, count >= 4).  sweep_total (nullable) accumulates the sweeps executed.
//   pass 2  after wn_launch_warm_transform_f64(..., uprev): every point started from uprev[k], the
//           eigenvectors of the same point of the previous mesh plane.
// ctr_stride: 0 updates sweeps[0] / sweep_total[0]; WN_CTR_STRIDE spreads the updates over the counter
// buckets of the status words (sweeps = words + WN_CTR_BASE, sweep_total = words + WN_CTR_BASE + 2).
hipError_t wn_launch_eig_pass_f64(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                                  int count, const WnShape &sh, int tpb, int pass, hipStream_t st,
                                  double2 *vscratch = nullptr, const double2 *uprev = nullptr, int ctr_stride = 0,
                                  int *status = nullptr);
// Register-resident-V Jacobi (wnlqg_eig.hip) for n <= 32: the same passes and stopping rule as the LDS
// kernel, two waves per workgroup (H in LDS, V rows in registers), 64 / N points per workgroup.  It is the
// default of wn_launch_eig_pass_f64 for n <= 32 (tpb 0 or WN_EIG_RV); tpb 64, 128, 256 select the LDS kernel.
#define WN_EIG_RV 1
// Deferred-eigenvector Jacobi (wnlqg_eig.hip) for n <= 32, the default there (tpb 0 or WN_EIG_DV): H-only sweeps
// at one wave per point (single LDS buffer, rotations recorded), then the recorded rounds applied to register
// rows of V; points needing more than the recorded sweeps are finished by the register-V kernel's roles.
// scr: wn_eigd_scratch_elems(n, count) complex elements.
#define WN_EIG_DV 2
size_t wn_eigd_scratch_elems(int n, int count);
hipError_t wn_launch_eig_defer(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                               int count, const WnShape &sh, int pass, hipStream_t st, const double2 *uprev,
                               int ctr_stride, int *status, double2 *scr);
bool wn_eig_rv_supported(int n);
hipError_t wn_launch_eig_rv(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                            int count, const WnShape &sh, int pass, hipStream_t st, const double2 *uprev,
                            int ctr_stride, int *status);
// size in complex elements of the eigensolver / rotation scratch for count points (n <= 32: the deferred-
// eigenvector records; 32 < n <= 64: 0)
// 64 < n <= 128 (wnlqg_eigbig.hip): sweep chunks on H with recorded rotations, rotations applied to V by
// row-owning waves, then sort and phase; scratch: wn_eig_big_scratch_elems(n, count) complex elements.
size_t wn_eig_big_scratch_elems(int n, int count);
hipError_t wn_launch_eig_big(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                             int count, const WnShape &sh, hipStream_t st, double2 *scr, int ctr_stride, int *status,
                             int pass = -1, const double2 *uprev = nullptr);
size_t wn_eig_scratch_elems(int n, int count);
size_t wn_rotate_scratch_elems(int n, int count);
bool wn_warm_start(int n, int count);
// vscratch: the eigensolver scratch, required for n > 64 (the transform stages its products there)
hipError_t wn_launch_warm_transform_f64(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                                        const double2 *uprev = nullptr, double2 *vscratch = nullptr);
hipError_t wn_launch_warm_big(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                              const double2 *uprev, double2 *scr);

// Warm-started mesh eigensolver groups: consecutive k points (neighbours along the fastest mesh axis) form
// groups of WN_WARM_G; the anchor of a group (position WN_WARM_A, or the last point of a short group) is solved
// from the identity (pass 0); every other point of the group is then solved from the anchor's eigenvectors
// (pass 1).  wn_warm_k maps the workgroup (or slot) b of a pass to its k point and, in pass 1, its anchor.
#define WN_WARM_G 4
#define WN_WARM_A 1
__device__ __forceinline__ int wn_warm_k(int b, int pass, int count, int *anchor) {
    if (pass == 0) {
        const int base = b * WN_WARM_G;
        *anchor = -1;
        return base + min(WN_WARM_A, count - 1 - base);
    }
    const int g = b / (WN_WARM_G - 1), j = b - g * (WN_WARM_G - 1);
    const int base = g * WN_WARM_G;
    const int al = min(WN_WARM_A, count - 1 - base);
    *anchor = base + al;
    return base + (j < al ? j : j + 1);
}
// Clears that are safe under stream capture: p[0 .. n) = 0; flags = {0, 0, 0, 0x7f7f7f7f}.
hipError_t wn_launch_zero(double *p, long long n, hipStream_t st);
// status words: [0] pair-gap flag, [1] eigensolver sweep cap reached, [2] geometry flags, [3] first failing k; then WN_CTR_BUCKETS counter
// buckets of WN_CTR_STRIDE words (128 bytes apart, so that the per-workgroup atomics of a launch spread
// over many addresses): bucket word [0] max sweeps, [2..3] total sweeps (64-bit), [4..5] total active
// SC/IC pairs (64-bit), [6] SC/IC points rotated to pair elements, [7] SC/IC points rotated in full,
// [8] bands of the element points, [9] element points whose dH and A were formed as rows (Hermitian model);
// workgroup g adds to bucket g % WN_CTR_BUCKETS.
#define WN_CTR_BASE 8
#define WN_CTR_STRIDE 32
#define WN_CTR_BUCKETS 64
#define WN_FLAG_WORDS (WN_CTR_BASE + WN_CTR_BUCKETS * WN_CTR_STRIDE)
hipError_t wn_launch_reset_flags(int *f, hipStream_t st);
// U^+ X U for channels 1 .. C-1, in place.
// scratch: wn_rotate_scratch_elems(n, count) complex elements, required for n > 64
hipError_t wn_launch_rotate_f64(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                                double2 *scratch = nullptr);
// Screening, pair lists, generalized derivatives and SC/IC kernels.
// (k_screen: occupations, windows and the active pair list of every k point, before the rotation.)
hipError_t wn_launch_screen(const double *e, const double *omega, int *count, int4 *pmeta, double2 *pmd, int *flags,
                            int nk, const WnShape &sh, const WnScreen &sc, hipStream_t st);
// SC/IC rotation: points without pairs skipped; points with at most WN_PCAP pairs get dH and A rotated in
// full and the pairs' dd H / dA elements in pwd[k][pair][direction][15] (wn_rotate_sc_elements true);
// otherwise, or when wn_rotate_sc_elements is false, every channel is rotated (wn_launch_rotate_f64).
#define WN_PCAP 8
bool wn_rotate_sc_elements(const WnShape &sh);
// ch32: channels 1 .. C-1 stored in FP32 by the mixed form's contraction (requires wn_rotate_sc_elements)
hipError_t wn_launch_rotate_sc(double *x, const double2 *u, const int *count, const int4 *pmeta, double2 *pwd,
                               int nk, const WnShape &sh, hipStream_t st, double2 *scratch, int *flags,
                               bool ch32 = false);
// Generalized derivatives and SC/IC kernels of the screened pairs (pwd: nullptr to read every element from
// the rotated channels).
hipError_t wn_launch_pairs_f64(const double *x, const double *e, const int *count, const int4 *pmeta,
                               const double2 *pwd, double2 *q, int nk, const WnShape &sh, const WnScreen &sc,
                               hipStream_t st);
// Per-point weights of a point-list integral: the occupation difference of every active pair of point k
// (pmd[k n^2 + j].y, j < count[k]) times w[k], so that the frequency contraction sums w_k value_k.
hipError_t wn_launch_weight_pairs(double2 *pmd, const int *count, const double *w, int nk, const WnShape &sh,
                                  hipStream_t st);
// Frequency contraction of the pairs of each chunk (chunks of at most WN_ACC_KMAX k points).
#define WN_ACC_KMAX 64
hipError_t wn_launch_accum(const int4 *chunks, int nchunks, const int *count, const int4 *pmeta, const double2 *pmd,
                           const double2 *q, const double *omega, double2 *partial, const WnShape &sh,
                           const WnScreen &sc, hipStream_t st);
// BCD moments or AHC curvature per chunk of k points: partial[chunk][comp][mu] (comp = a + 3b (+ 9c)),
// each point's value times wts[k] when wts is not null (weighted point lists).
// flags[2] |= 1 and flags[3] = min global k index on a spectral isolation failure; flags[2] |= 2 on a
// non-finite value.
hipError_t wn_launch_geom_f64(const double *x, const double *e, const double *mu, const int4 *chunks, int nchunks,
                              const double *wts, double2 *partial, int *flags, long long kbase, const WnShape &sh,
                              const WnGeom &gp, hipStream_t st);
// SHG: optical position derivatives der[k][0..8][n][n] and the operands of their commutator products
// (der holds WN_SHG_DER_BLOCKS n^2 complex per k point), pair lists and coefficient rows for `count_k` k
// points, then the frequency contraction of `nchunks` chunks into partial[chunk] = [19 nc][nw] complex
// followed by [nc] (one-band BCD, semiclassical) sums.
#define WN_SHG_DER_BLOCKS 21
hipError_t wn_launch_shg(const double *x, const double *e, const double *omega, const int4 *chunks, int nchunks,
                         int count_k, double2 *der, int *count, int *pidx, double *pdiff, double2 *coef,
                         double2 *one, double2 *partial, const WnShape &sh, const WnShg &sp, hipStream_t st);
// Fixed-order chunk -> lane sums.
hipError_t wn_launch_lane_reduce(const int4 *lanes, int nlanes, const double2 *partial, double2 *lanebuf,
                                 double2 *lanecomp, int nelem, hipStream_t st);
