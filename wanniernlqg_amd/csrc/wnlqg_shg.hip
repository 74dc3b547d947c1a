// Second-harmonic generation (SHG) per-k stage (reference
// src/Responses/SecondHarmonicGeneration/SecondHarmonicKernels.jl:71-294 and
// src/MatrixElements/Interpolated/OpticalPositionDerivatives.jl:13-143).
//
// Per k point (Hamiltonian-gauge channels from the rotation kernel, ascending eigenvalues):
//   occupations f_n (strict step at T <= 1e-7 K), active pairs n != m with |f_n - f_m| >= 1e-10;
//   bare and regularized vertices  r0 = A + i V/(E_m - E_n) (1/d cut below 1e-7 eV),
//   r_eta = A + i V (E_m - E_n)/((E_m - E_n)^2 + eta^2);
//   k_shg_der   : the optical position derivative der[a][b][n][m] of every active pair;
//   k_shg_coef  : per active pair and component (a,b,c) the seven-term coefficients of the three
//                 resonances d1 = R(dE - w), d2 = R(dE - 2w), dn = R(-dE - w), R(x) = PV(x) + i pi delta(x),
//                 and of the k-independent low-frequency factor; per band the one-band (Fermi-surface)
//                 BCD and semiclassical sums at T > 0;
//   k_shg_acc   : sum over the chunk's k points and pairs of resonance x coefficient for every
//                 frequency, in a fixed order, as real FP64 MFMA GEMMs whose A operand (the resonances)
//                 is generated in registers; k_shg_low: the frequency-independent block and one-band sums.
// The remaining factors depend on the frequency only (low = PV(w, w_low), the one-band kernel
// low - i delta(w, w_low)) and are applied to the lane sums on the host, which is exact because the
// k sum is linear.
//
// Coefficient columns of one pair (nc requested components, column = kind block + comp * width + j):
//   d1 block (width 8): t0q0 f inv second, t1q0 f inv third, t3q0 -f inv fourth/2, t4q0 -f fifth/4 (x low^2),
//                       t0q1 o second,    t1q1 o third,     t3q1 -o fourth/2,     t4q1 -o fifth/4 (x low)
//   d2 block (width 6): t0q0 2 f inv first, t1q0 -4 f inv third, t2q0 o tf, t0q1 o first, t1q1 -2 o third,
//                       t2q1 -i o dE tf
//   dn block (width 2): t2q0 o to/2, t2q1 i o dE to
//   low block (width 3): t0q0 f inv (first + second), t1q0 -f inv third, t3q0 -f inv fourth/2
// with o = f_n - f_m, f = i o / 2, inv = 1/dE, dE = E_m - E_n.
#include "wnlqg_common.h"
#include "wnlqg_kernels.h"

#include <math.h>

#include <algorithm>

#define WN_SHG_TPB 256
#define WN_SHG_NMAX 128   // orbitals (<= WN_SHG_TPB: one thread per band in the per-band phases)

namespace {

constexpr double kShgKb = 8.617333262145e-5;

__device__ inline double shg_occupation(double e, double mu, double temperature) {
    if (temperature <= 1.0e-7) return e <= mu ? 1.0 : 0.0;
    const double x = (e - mu) / (kShgKb * temperature);
    return x >= 0.0 ? exp(-x) / (1.0 + exp(-x)) : 1.0 / (1.0 + exp(x));
}

// 1/d with the 1e-7 eV cut (PositionVertexRegularization.jl:1-8)
__device__ inline double shg_bare_inverse(double d) { return fabs(d) < 1.0e-7 ? 0.0 : 1.0 / d; }

__device__ inline double shg_principal(double x, double w) { return x / (x * x + w * w); }

__device__ inline double shg_weight(double d, double eta) { return (eta * eta) / (d * d + eta * eta); }

// Channels of one k point (channel c at xk + c n^2), addressed by arithmetic rather than a pointer
// table, so that runtime component indices do not force the table into scratch memory.
struct ShgView {
    const double2 *xk;
    const double *E;
    int n, n2;
    int pos;               // position operator present (A and dA channels)
    double eta;
    __device__ __forceinline__ const double2 *A(int a) const { return xk + (size_t)(WN_CH_A + a) * n2; }
    __device__ __forceinline__ const double2 *V(int a) const { return xk + (size_t)(WN_CH_DH + a) * n2; }
    __device__ __forceinline__ const double2 *dA(int i) const { return xk + (size_t)(WN_CH_DA + i) * n2; }  // d_b A_a: 3a + b
    __device__ __forceinline__ const double2 *W(int s) const { return xk + (size_t)(WN_CH_DDH + s) * n2; }  // xx yy zz yz xz xy
};

__device__ inline double2 sv_A(const ShgView &s, int a, int i, int j) {
    return s.pos ? s.A(a)[i * s.n + j] : make_double2(0.0, 0.0);
}
__device__ inline double2 sv_V(const ShgView &s, int a, int i, int j) { return s.V(a)[i * s.n + j]; }
// bare vertex part V_ij / (E_j - E_i) (cut) and regularized part
__device__ inline double2 sv_bare(const ShgView &s, int a, int i, int j) {
    return cscale(sv_V(s, a, i, j), shg_bare_inverse(s.E[j] - s.E[i]));
}
__device__ inline double2 sv_reg(const ShgView &s, int a, int i, int j) {
    return cscale(sv_V(s, a, i, j), wn_reg_inverse(s.E[j] - s.E[i], s.eta));
}
// r0 = A + i bare, r_eta = A + i reg
__device__ inline double2 sv_dip(const ShgView &s, int a, int i, int j) {
    return cadd(sv_A(s, a, i, j), cmuli(sv_bare(s, a, i, j)));
}
__device__ inline double2 sv_rdip(const ShgView &s, int a, int i, int j) {
    return cadd(sv_A(s, a, i, j), cmuli(sv_reg(s, a, i, j)));
}

__device__ inline ShgView make_view(const double *x, const double *e, long long k, const WnShape &sh, double eta) {
    ShgView s;
    const int n = sh.n;
    s.xk = (const double2 *)(x + (size_t)k * sh.ldo);
    s.E = e + (size_t)k * n;
    s.n = n;
    s.n2 = n * n;
    s.pos = sh.has_pos;
    s.eta = eta;
    return s;
}

__device__ inline bool shg_active(const double *occ, int i, int j) { return i != j && fabs(occ[i] - occ[j]) >= 1.0e-10; }

}  // namespace

// ------------------------------------------------------------------------------------------------
// Optical position derivative (OpticalPositionDerivatives.jl:13-130) as commutator GEMMs.
// Per k point the sums over the intermediate band p of the reference are matrix products:
//   F_ab = [A_a, R_b]                                     (sum_a of the reference, before its diagonal term)
//   G_ab = i [V_a, R_b] + [A'_a, X_b] - [Q'_a, Y_b]        (i sum_h + the eta-correction numerator)
// with R_b = V_b o r(E_j - E_i) the regularized vertex, X_b = w o V_b, Y_b = w o A_b (w_ij = eta^2 /
// ((E_i - E_j)^2 + eta^2)), Q'_a = V_a + i (E_i - E_j) o A_a, and A'_a = A_a; X, Y, Q', A' with a zero
// diagonal, which removes exactly the reference's p = n and p = m terms of the eta correction (each of
// its four products has the index n on one factor's diagonal and m on the other's).  At T = 0 the
// active pairs are (occupied, empty) and (empty, occupied), and the occupied bands are a prefix of the
// ascending spectrum: only the two off-diagonal blocks rows [0, n_occ) x columns [n_occ, n) and rows
// [n_occ, n) x columns [0, n_occ) are computed (no Hermitian symmetry is assumed: the channels of a
// model whose H(-R) is not H(R)^+ are not Hermitian, and the reference takes them as they are).  At T > 0
// every element is computed.
// Per k point the derivative buffer holds WN_SHG_DER_BLOCKS n^2 blocks: der[3a+b] (9), R_b (3), X_b (3),
// Y_b (3), Q'_a (3).
// ------------------------------------------------------------------------------------------------
namespace {
constexpr int kShgR = 9, kShgX = 12, kShgY = 15, kShgQ = 18;   // first block of each operand family

// element (i, j) of the T = 0 block geometry: occupied bands (E <= mu) of one k point
__device__ inline int shg_occupied(const double *E, int n, double mu) {
    int c = 0;
    for (int i = 0; i < n; ++i) c += E[i] <= mu;
    return c;
}

// largest tile count (16 x 16 tiles) of one derivative matrix: both off-diagonal blocks at T = 0, the
// whole matrix at T > 0
int shg_vtx_tiles(int n, bool block) {
    const int t = (n + 15) / 16;
    if (!block) return t * t;
    int best = 0;
    for (int no = 1; no < n; ++no) best = std::max(best, 2 * ((no + 15) / 16) * ((n - no + 15) / 16));
    return std::max(best, 1);
}
}  // namespace

// Derived operands R_b, X_b, Y_b, Q'_a of one k point.  Grid (count, 12): block (k, operand).
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_prep(const double *__restrict__ x, const double *__restrict__ eg,
                                                        double2 *__restrict__ der, WnShape sh, WnShg sp) {
    const long long k = blockIdx.x;
    const int op = blockIdx.y, fam = op / 3, d = op - (op / 3) * 3;
    const int n = sh.n, n2 = n * n;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    double2 *out = der + ((size_t)k * WN_SHG_DER_BLOCKS + kShgR + op) * n2;
    for (int e = threadIdx.x; e < n2; e += WN_SHG_TPB) {
        const int i = e / n, j = e - (e / n) * n;
        double2 v;
        if (fam == 0) {
            v = sv_reg(s, d, i, j);
        } else if (i == j) {
            v = make_double2(0.0, 0.0);
        } else if (fam == 3) {
            v = cadd(sv_V(s, d, i, j), cmuli(cscale(sv_A(s, d, i, j), s.E[i] - s.E[j])));
        } else {
            const double w = shg_weight(s.E[i] - s.E[j], sp.eta);
            v = cscale(fam == 1 ? sv_V(s, d, i, j) : sv_A(s, d, i, j), w);
        }
        out[e] = v;
    }
}

// One block of a tile's product sum_p (f_L L[row][p]) (f_R R[p][col]) into the three Gauss accumulators of
// the tile (t1 = sum Re Re, t2 = sum Im Im, t3 = sum (Re + Im)(Re + Im)).  FL: 0 -> 1, 1 -> -1, 2 -> i;
// FR: 0 -> 1, 1 -> -1, 3 -> -i; ZL / ZR zero L's / R's diagonal (p == row / p == col).  Lane (lr, lk) takes
// K index p0 + lk of each step; four steps per iteration, the next iteration's operands loaded first; full
// iterations need no bounds (the tail iteration masks p >= n).
template <int FL, int FR, bool ZL, bool ZR>
__device__ __forceinline__ void vtx_block(const double2 *__restrict__ L, const double2 *__restrict__ R, int n, int row,
                                          int col, int lk, f64x4 &t1, f64x4 &t2, f64x4 &t3) {
    auto step = [&](double2 av, double2 cv, int p, bool valid) {
        double ar = FL == 2 ? -av.y : (FL == 1 ? -av.x : av.x);
        double ai = FL == 2 ? av.x : (FL == 1 ? -av.y : av.y);
        double cr = FR == 3 ? cv.y : (FR == 1 ? -cv.x : cv.x);
        double ci = FR == 3 ? -cv.x : (FR == 1 ? -cv.y : cv.y);
        if (!valid || (ZL && p == row)) ar = ai = 0.0;
        if (!valid || (ZR && p == col)) cr = ci = 0.0;
        t1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, cr, t1, 0, 0, 0);
        t2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, ci, t2, 0, 0, 0);
        t3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, cr + ci, t3, 0, 0, 0);
    };
    const double2 *lp = L + row * n + lk;    // L[row][p0 + lk]
    const double2 *rp = R + lk * n + col;    // R[p0 + lk][col]
    const int nfull = n & ~15;
    double2 lv[4], rv[4];
    if (nfull > 0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            lv[u] = lp[4 * u];
            rv[u] = rp[4 * u * n];
        }
    }
    for (int p0 = 0; p0 < nfull; p0 += 16) {
        double2 ln[4], rn[4];
        if (p0 + 16 < nfull) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                ln[u] = lp[p0 + 16 + 4 * u];
                rn[u] = rp[(p0 + 16 + 4 * u) * n];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) step(lv[u], rv[u], p0 + 4 * u + lk, true);
        if (p0 + 16 < nfull) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                lv[u] = ln[u];
                rv[u] = rn[u];
            }
        }
    }
    if (nfull < n) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = nfull + 4 * u + lk, pc = min(p, n - 1);
            step(L[row * n + pc], R[pc * n + col], p, p < n);
        }
    }
}

// One 16 x 16 tile of F_ab and G_ab per wave on v_mfma_f64_16x16x4f64, complex products as three real
// products (t1 = Lr Rr, t2 = Li Ri, t3 = (Lr + Li)(Rr + Ri); C = (t1 - t2, t3 - t1 - t2)), then the
// derivative of every element of the tile.  Lane l holds the left operand (row l%16, band p0 + l/16) and
// the right operand (band p0 + l/16, column l%16).  Tasks of one k point run on one XCD (workgroup b on
// XCD b % 8), so that its operands are read from that XCD's L2.
__global__ __launch_bounds__(64) void k_shg_vtx(const double *__restrict__ x, const double *__restrict__ eg,
                                               double2 *__restrict__ der, int count_k, int tiles, WnShape sh,
                                               WnShg sp) {
    const int tpk = 9 * tiles;
    const int b = blockIdx.x, xcd = b & 7, s8 = b >> 3;
    const long long k = (long long)(s8 / tpk) * 8 + xcd;
    if (k >= count_k) return;
    const int task = s8 - (s8 / tpk) * tpk;
    const int ab = task / tiles, tile = task - (task / tiles) * tiles;
    const int a = ab / 3, bb = ab - (ab / 3) * 3;
    const int n = sh.n, n2 = n * n;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    int r_lo = 0, r_hi = n, c_lo = 0, c_hi = n, t = tile;
    if (sp.temperature <= 1.0e-7) {
        // block (occupied rows, empty columns), then block (empty rows, occupied columns)
        const int no = shg_occupied(s.E, n, sp.mu);
        const int t1 = ((no + 15) / 16) * ((n - no + 15) / 16);
        if (t < t1) {
            r_hi = no;
            c_lo = no;
        } else {
            t -= t1;
            r_lo = no;
            c_hi = no;
        }
    }
    const int ntc = (c_hi - c_lo + 15) / 16, ntr = (r_hi - r_lo + 15) / 16;
    if (t >= ntr * ntc) return;
    const int ti = t / ntc, tj = t - (t / ntc) * ntc;
    const int r0 = r_lo + 16 * ti, c0 = c_lo + 16 * tj;
    const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
    const int row = min(r0 + lr, n - 1), col = min(c0 + lr, n - 1);   // operand element (clamped in bounds)
    const double2 *dk = der + (size_t)k * WN_SHG_DER_BLOCKS * n2;
    const bool pos = s.pos != 0, eta = pos && sp.eta_correction;
    f64x4 f1 = {0.0, 0.0, 0.0, 0.0}, f2 = f1, f3 = f1, g1 = f1, g2 = f1, g3 = f1;
    // products L R of the eight blocks; each block's factor on L or R (1, -1, i, -i) and zero diagonal are
    // compile-time (vtx_block), so applying them costs at most a swap and sign changes per element
    const double2 *Aa = s.A(a), *Va = s.V(a), *Rb = dk + (size_t)(kShgR + bb) * n2;
    const double2 *Xb = dk + (size_t)(kShgX + bb) * n2, *Yb = dk + (size_t)(kShgY + bb) * n2;
    const double2 *Qa = dk + (size_t)(kShgQ + a) * n2;
    if (pos) {
        vtx_block<0, 0, false, false>(Aa, Rb, n, row, col, lk, f1, f2, f3);    // F: A_a R_b
        vtx_block<0, 1, false, false>(Rb, Aa, n, row, col, lk, f1, f2, f3);    //    - R_b A_a
    }
    vtx_block<2, 0, false, false>(Va, Rb, n, row, col, lk, g1, g2, g3);        // G: i V_a R_b
    vtx_block<0, 3, false, false>(Rb, Va, n, row, col, lk, g1, g2, g3);        //    - i R_b V_a
    if (eta) {
        vtx_block<0, 0, true, false>(Aa, Xb, n, row, col, lk, g1, g2, g3);     //    A'_a X_b
        vtx_block<0, 1, false, true>(Xb, Aa, n, row, col, lk, g1, g2, g3);     //    - X_b A'_a
        vtx_block<1, 0, false, false>(Qa, Yb, n, row, col, lk, g1, g2, g3);    //    - Q'_a Y_b
        vtx_block<0, 0, false, false>(Yb, Qa, n, row, col, lk, g1, g2, g3);    //    Y_b Q'_a
    }
    // element (nn, m) = (r0 + lk + 4q, c0 + lr) of the tile
    const int m = c0 + lr;
    double2 *outb = der + ((size_t)k * WN_SHG_DER_BLOCKS + ab) * n2;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int nn = r0 + lk + 4 * q;
        if (nn >= r_hi || m >= c_hi) continue;
        const double2 ca = make_double2(f1[q] - f2[q], f3[q] - f1[q] - f2[q]);
        const double2 gg = make_double2(g1[q] - g2[q], g3[q] - g1[q] - g2[q]);
        const double2 Aan = sv_A(s, a, nn, nn), Aam = sv_A(s, a, m, m);
        const double2 Abn = sv_A(s, bb, nn, nn), Abm = sv_A(s, bb, m, m);
        const double2 Van = sv_V(s, a, nn, nn), Vam = sv_V(s, a, m, m);
        const double2 Vbn = sv_V(s, bb, nn, nn), Vbm = sv_V(s, bb, m, m);
        const double2 reg_b = dk[(size_t)(kShgR + bb) * n2 + nn * n + m];
        const double2 bare_a = sv_bare(s, a, nn, m), bare_b = sv_bare(s, bb, nn, m);
        const double2 dAa = csub(Aan, Aam), dAb = csub(Abn, Abm);
        const double2 sum_a = csub(ca, cmul(dAa, reg_b));
        const double2 dAab = pos ? s.dA(3 * a + bb)[nn * n + m] : make_double2(0.0, 0.0);
        double2 finite = cadd(dAab, sum_a);
        finite = cadd(finite, cmul(dAa, bare_b));
        finite = cadd(finite, cmul(dAb, bare_a));
        finite = csub(finite, cmuli(cmul(sv_A(s, a, nn, m), dAb)));
        // numerator = i (W_ab + [V_a, R_b] - dV_a reg_b + bare_a dV_b + bare_b dV_a) + eta numerator
        const double2 Wab = s.W(wn_sym(a, bb))[nn * n + m];
        double2 inner = csub(Wab, cmul(csub(Van, Vam), reg_b));
        inner = cadd(inner, cscale(bare_a, Vbn.x - Vbm.x));
        inner = cadd(inner, cscale(bare_b, Van.x - Vam.x));
        const double2 numerator = cadd(cmuli(inner), gg);
        const double iv = shg_bare_inverse(s.E[m] - s.E[nn]);
        const double2 v = iv == 0.0 ? finite : cadd(finite, cscale(numerator, iv));
        outb[nn * n + m] = v;
    }
}

// ------------------------------------------------------------------------------------------------
// Pair list of one k point (one workgroup per k): active pairs n != m with |f_n - f_m| >= 1e-10 in
// n-major order (SecondHarmonicKernels.jl:140-141): count[k], pdiff[k][p] = E_m - E_n and the index map
// pidx[k][n][m] = p (-1 for an inactive pair); and at T > 0 the one-band (Fermi-surface) BCD and
// semiclassical sums one[k][comp] = (bcd, semi) (SecondHarmonicKernels.jl:245-294).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_pairs(const double *__restrict__ x, const double *__restrict__ eg,
                                                         int *__restrict__ count, int *__restrict__ pidx,
                                                         double *__restrict__ pdiff, double2 *__restrict__ one,
                                                         WnShape sh, WnShg sp) {
    __shared__ double occ[WN_SHG_NMAX];
    __shared__ int rowcnt[WN_SHG_NMAX];
    __shared__ int rowoff[WN_SHG_NMAX + 1];
    extern __shared__ double lds_one[];  // [n][nc][2] one-band contributions
    const long long k = blockIdx.x;
    const int n = sh.n, n2 = n * n, nc = sp.ncomp;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    const int t = threadIdx.x;
    if (t < n) occ[t] = shg_occupation(s.E[t], sp.mu, sp.temperature);
    __syncthreads();
    if (t < n) {
        int c = 0;
        for (int m = 0; m < n; ++m) c += shg_active(occ, t, m);
        rowcnt[t] = c;
    }
    __syncthreads();
    if (t == 0) {
        rowoff[0] = 0;
        for (int i = 0; i < n; ++i) rowoff[i + 1] = rowoff[i] + rowcnt[i];
        count[k] = rowoff[n];
    }
    __syncthreads();
    const size_t pbase = (size_t)k * n2;
    if (t < n) {
        int p = rowoff[t];
        for (int m = 0; m < n; ++m) {
            if (shg_active(occ, t, m)) {
                pidx[pbase + t * n + m] = p;
                pdiff[pbase + p] = s.E[m] - s.E[t];
                ++p;
            } else {
                pidx[pbase + t * n + m] = -1;
            }
        }
    }
    double2 *ok = one + (size_t)k * nc;
    if (sp.temperature <= 1.0e-7) {
        for (int ci = t; ci < nc; ci += WN_SHG_TPB) ok[ci] = make_double2(0.0, 0.0);
        return;
    }
    const double kt = kShgKb * sp.temperature;
    for (int item = t; item < n * nc; item += WN_SHG_TPB) {
        const int nn = item / nc, ci = item - (item / nc) * nc;
        double bcd = 0.0, semi = 0.0;
        const double off = s.E[nn] - sp.mu;
        if (fabs(off) <= 5.0 * kt) {
            const double d1 = -exp(-0.5 * (off / kt) * (off / kt)) / (kt * sqrt(2.0 * M_PI));
            const double d2 = -off / (kt * kt) * d1;
            const bool isolated = (nn == 0 || s.E[nn] - s.E[nn - 1] > sp.tol) && (nn == n - 1 || s.E[nn + 1] - s.E[nn] > sp.tol);
            const int cc = sp.comp[ci];
            const int a = cc % 3, b = (cc / 3) % 3, c = cc / 9;
            double hess = s.W(wn_sym(b, c))[nn * n + nn].x;
            for (int q = 0; q < n; ++q) {
                if (q == nn) continue;
                const double2 pr = cmul(sv_V(s, b, nn, q), sv_V(s, c, q, nn));
                hess += 2.0 * pr.x * shg_principal(s.E[nn] - s.E[q], sp.eta);
            }
            const double va = sv_V(s, a, nn, nn).x, vb = sv_V(s, b, nn, nn).x, vc = sv_V(s, c, nn, nn).x;
            semi = d2 * va * vb * vc + d1 * va * hess;
            if (isolated) {
                // curv[a][y] for y = c and y = b (zero on the diagonal)
                double cv[2] = {0.0, 0.0};
                const int ys[2] = {c, b};
                for (int u = 0; u < 2; ++u) {
                    const int y = ys[u];
                    if (y == a) continue;
                    double2 val = s.pos ? csub(s.dA(3 * y + a)[nn * n + nn], s.dA(3 * a + y)[nn * n + nn])
                                          : make_double2(0.0, 0.0);
                    for (int q = 0; q < n; ++q) {
                        const double2 ba_qn = sv_bare(s, a, q, nn), ba_nq = sv_bare(s, a, nn, q);
                        const double2 by_qn = sv_bare(s, y, q, nn), by_nq = sv_bare(s, y, nn, q);
                        val = cadd(val, cmul(sv_A(s, y, nn, q), ba_qn));
                        val = csub(val, cmul(ba_nq, sv_A(s, y, q, nn)));
                        val = csub(val, cmul(sv_A(s, a, nn, q), by_qn));
                        val = cadd(val, cmul(by_nq, sv_A(s, a, q, nn)));
                        val = csub(val, cmuli(csub(cmul(ba_nq, by_qn), cmul(by_nq, ba_qn))));
                    }
                    cv[u] = val.x;
                }
                bcd = d1 * (vb * cv[0] + vc * cv[1]);
            }
        }
        lds_one[(nn * nc + ci) * 2] = bcd;
        lds_one[(nn * nc + ci) * 2 + 1] = semi;
    }
    __syncthreads();
    for (int ci = t; ci < nc; ci += WN_SHG_TPB) {
        double sb = 0.0, ss = 0.0;
        for (int nn = 0; nn < n; ++nn) {
            sb += lds_one[(nn * nc + ci) * 2];
            ss += lds_one[(nn * nc + ci) * 2 + 1];
        }
        ok[ci] = make_double2(sb, ss);
    }
}

// ------------------------------------------------------------------------------------------------
// Coefficient rows of the active pairs (SecondHarmonicKernels.jl:115-236): one thread per pair (n, m) of a
// 16 x 16 tile over the same blocks as the vertex kernel.  The three-band sums over q != n, m of the
// reference,
//   tf = r0_a,nm sum_q PV(2E_q - E_n - E_m) (r_b,mq r_c,qn + r_c,mq r_b,qn)
//   to = sum_q PV(2E_n - E_m - E_q) r_a,mq (r_b,qn r0_c,nm + r_c,qn r0_b,nm)
//        + PV(2E_m - E_q - E_n) r_a,qn (r0_b,nm r_c,mq + r0_c,nm r_b,mq)        (r = r_eta = A + i R),
// are regrouped with P_xy = r_x,mq r_y,qn as tf = r0_a,nm T_bc, to = r0_c,nm Z_ab + r0_b,nm Z_ac, where
// T_xy = sum_q PV0 (P_xy + P_yx) and Z_xy = sum_q PV1 P_xy + PV2 P_yx: fifteen sums per pair serve every
// component, and the three principal values are evaluated once per (pair, q).  Rows m (r_x,m.) and
// columns n (r_y,.n) of the tile are staged through LDS in chunks of WN_SHG_QC bands.
// Coefficient row layout (complex columns of component ci): d1 block ci 8 .. ci 8 + 7, d2 block
// 8 nc + ci 6 .., dn block 14 nc + ci 2 .., low block 16 nc + ci 3 ..
// ------------------------------------------------------------------------------------------------
#define WN_SHG_QC 32
namespace {
template <int N>
__device__ __forceinline__ double2 shg_pick(const double2 (&v)[N], int idx) {
    unsigned long long re = 0, im = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const unsigned long long msk = 0ull - (unsigned long long)(i == idx);
        re |= (unsigned long long)__double_as_longlong(v[i].x) & msk;
        im |= (unsigned long long)__double_as_longlong(v[i].y) & msk;
    }
    return make_double2(__longlong_as_double((long long)re), __longlong_as_double((long long)im));
}
}  // namespace
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_coef(const double *__restrict__ x, const double *__restrict__ eg,
                                                        const double2 *__restrict__ der, const int *__restrict__ pidx,
                                                        double2 *__restrict__ coef, int count_k, int tiles,
                                                        WnShape sh, WnShg sp) {
    __shared__ double2 rm[3][WN_SHG_QC][16];   // r_x[m][q] of the tile's columns m
    __shared__ double2 cn[3][WN_SHG_QC][16];   // r_y[q][n] of the tile's rows n
    __shared__ double es[WN_SHG_NMAX];
    const int b8 = blockIdx.x, xcd = b8 & 7, s8 = b8 >> 3;
    const long long k = (long long)(s8 / tiles) * 8 + xcd;
    if (k >= count_k) return;
    int tile = s8 - (s8 / tiles) * tiles;
    const int n = sh.n, n2 = n * n, nc = sp.ncomp, ncol = 19 * nc;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    int r_lo = 0, r_hi = n, c_lo = 0, c_hi = n;
    if (sp.temperature <= 1.0e-7) {
        const int no = shg_occupied(s.E, n, sp.mu);
        const int t1 = ((no + 15) / 16) * ((n - no + 15) / 16);
        if (tile < t1) {
            r_hi = no;
            c_lo = no;
        } else {
            tile -= t1;
            r_lo = no;
            c_hi = no;
        }
    }
    const int ntc = (c_hi - c_lo + 15) / 16, ntr = (r_hi - r_lo + 15) / 16;
    if (tile >= ntr * ntc) return;   // uniform over the workgroup
    const int r0 = r_lo + 16 * (tile / ntc), c0 = c_lo + 16 * (tile - (tile / ntc) * ntc);
    const int t = threadIdx.x, ln = t >> 4, lm = t & 15;
    const int nn = r0 + ln, m = c0 + lm;
    const bool inside = nn < r_hi && m < c_hi;
    const double2 *dk = der + (size_t)k * WN_SHG_DER_BLOCKS * n2;
    const double2 *rk = dk + (size_t)kShgR * n2;
    auto RE = [&](int d, int i, int j) { return cadd(sv_A(s, d, i, j), cmuli(rk[(size_t)d * n2 + i * n + j])); };
    for (int i = t; i < n; i += WN_SHG_TPB) es[i] = s.E[i];
    double2 T[6], Z[9];
#pragma unroll
    for (int i = 0; i < 6; ++i) T[i] = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 9; ++i) Z[i] = make_double2(0.0, 0.0);
    const double eta = sp.eta;
    const double en = s.E[min(nn, n - 1)], em = s.E[min(m, n - 1)];
    for (int q0 = 0; q0 < n; q0 += WN_SHG_QC) {
        __syncthreads();
        for (int e = t; e < 3 * WN_SHG_QC * 16; e += WN_SHG_TPB) {
            const int d = e / (WN_SHG_QC * 16), rem = e - d * (WN_SHG_QC * 16);
            {   // row element: q fastest (contiguous in memory)
                const int q = rem % WN_SHG_QC, ml = rem / WN_SHG_QC;
                rm[d][q][ml] = RE(d, min(c0 + ml, n - 1), min(q0 + q, n - 1));
            }
            {   // column element: n fastest
                const int nl = rem & 15, q = rem >> 4;
                cn[d][q][nl] = RE(d, min(q0 + q, n - 1), min(r0 + nl, n - 1));
            }
        }
        __syncthreads();
        const int qn = min(WN_SHG_QC, n - q0);
        for (int q = 0; q < qn; ++q) {
            const int qq = q0 + q;
            const double eq = es[qq];
            const bool skip = qq == nn || qq == m;
            const double pv0 = skip ? 0.0 : shg_principal(2.0 * eq - en - em, eta);
            const double pv1 = skip ? 0.0 : shg_principal(2.0 * en - em - eq, eta);
            const double pv2 = skip ? 0.0 : shg_principal(2.0 * em - eq - en, eta);
            double2 rx[3], ry[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                rx[d] = rm[d][q][lm];
                ry[d] = cn[d][q][ln];
            }
            double2 P[3][3];
#pragma unroll
            for (int xd = 0; xd < 3; ++xd)
#pragma unroll
                for (int yd = 0; yd < 3; ++yd) P[xd][yd] = cmul(rx[xd], ry[yd]);
#pragma unroll
            for (int xd = 0; xd < 3; ++xd)
#pragma unroll
                for (int yd = 0; yd < 3; ++yd) {
                    Z[3 * xd + yd] = cadd(Z[3 * xd + yd], cadd(cscale(P[xd][yd], pv1), cscale(P[yd][xd], pv2)));
                    if (xd <= yd) {
                        const int si = wn_sym(xd, yd);
                        T[si] = cadd(T[si], cscale(cadd(P[xd][yd], P[yd][xd]), pv0));
                    }
                }
        }
    }
    if (!inside) return;
    const int p = pidx[(size_t)k * n2 + nn * n + m];
    if (p < 0) return;
    const double diff = s.E[m] - s.E[nn];
    const double inverse = shg_bare_inverse(diff);
    const double o = shg_occupation(s.E[nn], sp.mu, sp.temperature) - shg_occupation(s.E[m], sp.mu, sp.temperature);
    auto D = [&](int aa, int bb, int i, int j) { return dk[(size_t)(3 * aa + bb) * n2 + i * n + j]; };
    double2 *row = coef + ((size_t)k * n2 + p) * ncol;
    for (int ci = 0; ci < nc; ++ci) {
        const int cc = sp.comp[ci];
        const int a = cc % 3, b = (cc / 3) % 3, c = cc / 9;
        const double2 dip_a_nm = sv_dip(s, a, nn, m);
        const double2 dip_b_mn = sv_dip(s, b, m, nn), dip_c_mn = sv_dip(s, c, m, nn);
        const double2 dip_b_nm = sv_dip(s, b, nn, m), dip_c_nm = sv_dip(s, c, nn, m);
        const double2 first = cmul(dip_a_nm, cadd(D(b, c, m, nn), D(c, b, m, nn)));
        const double2 second = cadd(cmul(D(a, b, nn, m), dip_c_mn), cmul(D(a, c, nn, m), dip_b_mn));
        const double vc = sv_V(s, c, m, m).x - sv_V(s, c, nn, nn).x;
        const double vb = sv_V(s, b, m, m).x - sv_V(s, b, nn, nn).x;
        const double2 third =
            cscale(cmul(dip_a_nm, cadd(cscale(dip_b_mn, vc), cscale(dip_c_mn, vb))), inverse);
        const double2 fourth = cadd(cmul(D(b, a, nn, m), dip_c_mn), cmul(D(c, a, nn, m), dip_b_mn));
        const double va = sv_V(s, a, nn, nn).x - sv_V(s, a, m, m).x;
        const double2 fifth = cscale(cadd(cmul(dip_b_nm, dip_c_mn), cmul(dip_c_nm, dip_b_mn)), va);
        // the runtime component indices pick among the sums by bit masks (a dynamic register index, or
        // selects that the compiler turns into one, would place the sums in scratch memory)
        const double2 Tbc = shg_pick<6>(T, wn_sym(b, c)), Zab = shg_pick<9>(Z, 3 * a + b);
        const double2 Zac = shg_pick<9>(Z, 3 * a + c);
        const double2 tf = cmul(dip_a_nm, Tbc);
        const double2 to = cadd(cmul(dip_c_nm, Zab), cmul(dip_b_nm, Zac));
        // f = i o / 2;  f * inv * z = i (o inv / 2) z
        const double h = 0.5 * o;
        const double hi = h * inverse;
        double2 *d1 = row + ci * 8;
        double2 *d2 = row + nc * 8 + ci * 6;
        double2 *dn = row + nc * 14 + ci * 2;
        double2 *lw = row + nc * 16 + ci * 3;
        d1[0] = cmuli(cscale(second, hi));
        d1[1] = cmuli(cscale(third, hi));
        d1[2] = cmuli(cscale(fourth, -0.5 * hi));
        d1[3] = cmuli(cscale(fifth, -0.25 * h));
        d1[4] = cscale(second, o);
        d1[5] = cscale(third, o);
        d1[6] = cscale(fourth, -0.5 * o);
        d1[7] = cscale(fifth, -0.25 * o);
        d2[0] = cmuli(cscale(first, 2.0 * hi));
        d2[1] = cmuli(cscale(third, -4.0 * hi));
        d2[2] = cscale(tf, o);
        d2[3] = cscale(first, o);
        d2[4] = cscale(third, -2.0 * o);
        d2[5] = cmuli(cscale(tf, -o * diff));
        dn[0] = cscale(to, 0.5 * o);
        dn[1] = cmuli(cscale(to, o * diff));
        lw[0] = cmuli(cscale(cadd(first, second), hi));
        lw[1] = cmuli(cscale(third, -hi));
        lw[2] = cmuli(cscale(fourth, -0.5 * hi));
    }
}

// ------------------------------------------------------------------------------------------------
// Frequency contraction on FP64 MFMA.  For one chunk the partial sums of the three resonance kinds
// d1 = R(dE - w), d2 = R(dE - 2w), dn = R(-dE - w) form complex GEMMs out[w][c] = sum_p R_kind(dE_p, w)
// coef[p][c] over the kind's column block, written as real GEMMs with K = 2 x pairs.  Four pairs make
// two K steps of v_mfma_f64_16x16x4f64: lane l evaluates the complex resonance of frequency l%16 and
// pair l/16 of the group once and passes Re R as its A element of the first step and Im R in the
// second; its B elements are (Cr, Ci) of the pair's coefficient in the first step and (-Ci, Cr) in the
// second (real column 2c = Re, 2c+1 = Im of complex column c), both from one 16-byte load:
//   Re out = sum Re R Cr - Im R Ci,   Im out = sum Re R Ci + Im R Cr.
// A column tile is 16 real columns (8 complex) of one kind; the kinds' tiles are numbered d1 (nc tiles),
// d2 (ceil(6 nc / 8)), dn (ceil(2 nc / 8)).  One wave = one task: RTW row tiles of 16 frequencies times
// up to MAXT consecutive column tiles, over all pairs of its chunk in pair order (k points in order).
// Tasks of one chunk run on one XCD (workgroup b on XCD b % 8) so that its coefficient rows are read
// from that XCD's L2.  The frequency-independent low block and the one-band sums go to k_shg_low.
// ------------------------------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double2 shg_resonance(double x, double w2, double iw, double gnorm, double cut,
                                                 int gaussian) {
    // R(x) = PV(x) + i pi delta(x) (SecondHarmonicKernels.jl:86-98): PV = x / (x^2 + w^2); Gaussian
    // pi delta = sqrt(pi) exp(-(x/w)^2) / w (0 beyond 27.3 w, where exp underflows in FP64), Lorentzian
    // pi delta = w / (x^2 + w^2)
    const double pv = x * wn_rcp(x * x + w2);
    double d = 0.0;
    if (gaussian) {
        if (fabs(x) < cut) {
            const double y = x * iw;
            d = gnorm * exp(-y * y);
        }
    } else {
        const double y = x * iw;
        d = iw * wn_rcp(1.0 + y * y);
    }
    return make_double2(pv, d);
}

// pi delta(x) given 1 / (x^2 + w^2): Gaussian sqrt(pi) exp(-(x/w)^2) / w (0 beyond 27.3 w), Lorentzian w / (x^2 + w^2)
__device__ __forceinline__ double shg_delta(double x, double inv, double iw, double gnorm, double cut, double w,
                                            int gaussian) {
    if (!gaussian) return w * inv;
    double d = 0.0;
    if (fabs(x) < cut) {
        const double y = x * iw;
        d = gnorm * exp(-y * y);
    }
    return d;
}

struct ShgTileSet {
    int t0, t1, t2;   // tiles of d1, d2, dn
    __device__ __host__ static ShgTileSet of(int nc) { return {nc, (6 * nc + 7) / 8, (2 * nc + 7) / 8}; }
    __device__ __host__ int total() const { return t0 + t1 + t2; }
};

}  // namespace

template <int RTW, int N0, int N1, int N2>
__global__ __launch_bounds__(64, 2) void k_shg_acc(const int4 *__restrict__ chunks, int nchunks,
                                                const int *__restrict__ count, const double *__restrict__ pdiff,
                                                const double2 *__restrict__ coef, const double *__restrict__ omega,
                                                double2 *__restrict__ partial, int pcap, int nrg, int ncg, WnShg sp) {
    constexpr int NT = N0 + N1 + N2;
    const int ntask = nrg * ncg;
    const int b = blockIdx.x, xcd = b & 7, s = b >> 3;
    const int chunk = (s / ntask) * 8 + xcd;
    if (chunk >= nchunks) return;
    const int task = s - (s / ntask) * ntask;
    const int rg = task / ncg, cg = task - (task / ncg) * ncg;
    const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4, part = lr & 1;
    const int nc = sp.ncomp, ncol = 19 * nc, nw = sp.nw;
    // this lane's complex column of every tile of the task (B operand and output): tiles 0..N0-1 of kind
    // d1 from tile cg N0 of that kind, then N1 of d2 from cg N1, then N2 of dn from cg N2
    int col[NT];
    bool cv[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int kind = j < N0 ? 0 : (j < N0 + N1 ? 1 : 2);
        const int tt = kind == 0 ? cg * N0 + j : (kind == 1 ? cg * N1 + j - N0 : cg * N2 + j - N0 - N1);
        const int cbeg = kind == 0 ? 0 : (kind == 1 ? 8 * nc : 14 * nc);
        const int cw = kind == 0 ? 8 * nc : (kind == 1 ? 6 * nc : 2 * nc);
        const int cc = tt * 8 + (lr >> 1);
        cv[j] = cc < cw;
        col[j] = cv[j] ? cbeg + cc : 0;
    }
    double om[RTW];
    bool wv[RTW];
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
        const int w = (rg * RTW + r) * 16 + lr;
        wv[r] = w < nw;
        om[r] = wv[r] ? omega[w] : 0.0;
    }
    f64x4 acc[RTW][NT];
#pragma unroll
    for (int r = 0; r < RTW; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
    const double cut = 27.3 * sp.width;
    const double w2 = sp.width * sp.width, iw = 1.0 / sp.width;
    const double gnorm = sqrt(M_PI) * iw;
    const int4 ch = chunks[chunk];
    const int kend = ch.x + ch.y;
    // groups of four pairs in chunk order
    for (int k = ch.x; k < kend; ++k) {
        const int cnt = count[k];
        const double2 *ck = coef + (size_t)k * pcap * ncol;     // this point's rows (uniform base)
        const double *dk = pdiff + (size_t)k * pcap;
        for (int f0 = 0; f0 < cnt; f0 += 4) {
            const bool full = f0 + 4 <= cnt;
            const bool pv = f0 + lg < cnt;
            // a lane past the last pair reads the group's first pair (finite) and contributes R = 0; lanes of
            // frequencies past the axis feed only output rows that are not stored
            const int row = f0 + (pv ? lg : 0);
            const double dE = dk[row];
            double2 bv[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j) bv[j] = ck[row * ncol + col[j]];
            double2 R0[RTW], R1[RTW], R2[RTW];
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
                if (N0 && N1 && N2) {
                    // the three principal-value denominators x^2 + w^2 share one refined reciprocal
                    const double x0 = dE - om[r], x1 = dE - 2.0 * om[r], x2 = -dE - om[r];
                    const double a0 = x0 * x0 + w2, a1 = x1 * x1 + w2, a2 = x2 * x2 + w2;
                    const double a12 = a1 * a2;
                    const double q = wn_rcp(a0 * a12);
                    const double i0 = a12 * q, i1 = a0 * a2 * q, i2 = a0 * a1 * q;
                    R0[r] = make_double2(x0 * i0, shg_delta(x0, i0, iw, gnorm, cut, sp.width, sp.gaussian));
                    R1[r] = make_double2(x1 * i1, shg_delta(x1, i1, iw, gnorm, cut, sp.width, sp.gaussian));
                    R2[r] = make_double2(x2 * i2, shg_delta(x2, i2, iw, gnorm, cut, sp.width, sp.gaussian));
                } else {
                    if (N0) R0[r] = shg_resonance(dE - om[r], w2, iw, gnorm, cut, sp.gaussian);
                    if (N1) R1[r] = shg_resonance(dE - 2.0 * om[r], w2, iw, gnorm, cut, sp.gaussian);
                    if (N2) R2[r] = shg_resonance(-dE - om[r], w2, iw, gnorm, cut, sp.gaussian);
                }
                if (!full && !pv) {
                    R0[r] = R1[r] = R2[r] = make_double2(0.0, 0.0);
                }
            }
            // first K step: Re R against (Cr, Ci); second: Im R against (-Ci, Cr)
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double bA = cv[j] ? (part ? bv[j].y : bv[j].x) : 0.0;
#pragma unroll
                for (int r = 0; r < RTW; ++r) {
                    const double a = j < N0 ? R0[r].x : (j < N0 + N1 ? R1[r].x : R2[r].x);
                    acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bA, acc[r][j], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double bB = cv[j] ? (part ? bv[j].x : -bv[j].y) : 0.0;
#pragma unroll
                for (int r = 0; r < RTW; ++r) {
                    const double a = j < N0 ? R0[r].y : (j < N0 + N1 ? R1[r].y : R2[r].y);
                    acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bB, acc[r][j], 0, 0, 0);
                }
            }
        }
    }
    // D register q of lane (lg, lr) -> frequency row lg + 4q, real column lr of the tile
    double2 *o = partial + (size_t)chunk * ((size_t)ncol * nw + nc);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        if (cv[j]) {
#pragma unroll
            for (int r = 0; r < RTW; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int w = (rg * RTW + r) * 16 + lg + 4 * q;
                    if (w < nw) ((double *)(o + (size_t)col[j] * nw + w))[part] = acc[r][j][q];
                }
        }
    }
}

// Frequency-independent low block (R = 1) and the one-band sums of one chunk: thread t sums the pairs
// t, t + 256, ... of every column, then the 256 partial sums are combined in a fixed tree order; the low
// sums are stored for every frequency.
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_low(const int4 *__restrict__ chunks, const int *__restrict__ count,
                                                       const double2 *__restrict__ coef,
                                                       const double2 *__restrict__ one, double2 *__restrict__ partial,
                                                       int pcap, WnShg sp) {
    __shared__ double2 red[WN_SHG_TPB];
    const int nc = sp.ncomp, ncol = 19 * nc, nw = sp.nw;
    const int chunk = blockIdx.x, t = threadIdx.x;
    const int4 ch = chunks[chunk];
    double2 *o = partial + (size_t)chunk * ((size_t)ncol * nw + nc);
    for (int c = 0; c < 3 * nc; ++c) {
        double2 acc = make_double2(0.0, 0.0);
        for (int ki = 0; ki < ch.y; ++ki) {
            const int k = ch.x + ki, cnt = count[k];
            for (int f = t; f < cnt; f += WN_SHG_TPB)
                acc = cadd(acc, coef[((size_t)k * pcap + f) * ncol + 16 * nc + c]);
        }
        red[t] = acc;
        __syncthreads();
        for (int h = WN_SHG_TPB / 2; h > 0; h >>= 1) {
            if (t < h) red[t] = cadd(red[t], red[t + h]);
            __syncthreads();
        }
        const double2 sum = red[0];
        for (int w = t; w < nw; w += WN_SHG_TPB) o[(size_t)(16 * nc + c) * nw + w] = sum;
        __syncthreads();
    }
    if (t < nc) {
        double sb = 0.0, ss = 0.0;
        for (int k = ch.x; k < ch.x + ch.y; ++k) {
            const double2 v = one[(size_t)k * nc + t];
            sb += v.x;
            ss += v.y;
        }
        o[(size_t)ncol * nw + t] = make_double2(sb, ss);
    }
}

hipError_t wn_launch_shg(const double *x, const double *e, const double *omega, const int4 *chunks, int nchunks,
                         int count_k, double2 *der, int *count, int *pidx, double *pdiff, double2 *coef,
                         double2 *one, double2 *partial, const WnShape &sh, const WnShg &sp, hipStream_t st) {
    if (sh.n > WN_SHG_NMAX || sp.ncomp < 1 || sp.ncomp > 18) return hipErrorInvalidValue;
    if (count_k > 0) {
        hipLaunchKernelGGL(k_shg_prep, dim3(count_k, 12), dim3(WN_SHG_TPB), 0, st, x, e, der, sh, sp);
        const int tiles = shg_vtx_tiles(sh.n, sp.temperature <= 1.0e-7);
        const unsigned nb = (unsigned)((count_k + 7) / 8 * 8) * 9u * (unsigned)tiles;
        hipLaunchKernelGGL(k_shg_vtx, dim3(nb), dim3(64), 0, st, x, e, der, count_k, tiles, sh, sp);
        hipLaunchKernelGGL(k_shg_pairs, dim3(count_k), dim3(WN_SHG_TPB), (size_t)sh.n * sp.ncomp * 2 * sizeof(double),
                           st, x, e, count, pidx, pdiff, one, sh, sp);
        hipLaunchKernelGGL(k_shg_coef, dim3((unsigned)((count_k + 7) / 8 * 8) * (unsigned)tiles), dim3(WN_SHG_TPB), 0,
                           st, x, e, der, pidx, coef, count_k, tiles, sh, sp);
    }
    if (nchunks > 0) {
        // two row tiles of 16 frequencies per task (three spill at eight column tiles); nc <= 4: one task
        // covers every column tile of the three kinds; larger nc: tasks of eight tiles of one kind (tiles past
        // the kind's columns contribute zeros)
        const int nrt = (sp.nw + 15) / 16;
        const int nrg = (nrt + 1) / 2;
        const ShgTileSet ts = ShgTileSet::of(sp.ncomp);
        auto grid = [&](int ncg) { return dim3((unsigned)((nchunks + 7) / 8 * 8 * nrg * ncg)); };
#define WN_SHG_ACC(A, B, C, NCG)                                                                                    \
    hipLaunchKernelGGL((k_shg_acc<2, A, B, C>), grid(NCG), dim3(64), 0, st, chunks, nchunks, count, pdiff, coef, omega, \
                       partial, sh.n * sh.n, nrg, NCG, sp)
        switch (sp.ncomp) {
            case 1: WN_SHG_ACC(1, 1, 1, 1); break;
            case 2: WN_SHG_ACC(2, 2, 1, 1); break;
            case 3: WN_SHG_ACC(3, 3, 1, 1); break;
            case 4: WN_SHG_ACC(4, 3, 1, 1); break;
            default:
                WN_SHG_ACC(8, 0, 0, (ts.t0 + 7) / 8);
                WN_SHG_ACC(0, 8, 0, (ts.t1 + 7) / 8);
                WN_SHG_ACC(0, 0, 8, (ts.t2 + 7) / 8);
        }
#undef WN_SHG_ACC
        hipLaunchKernelGGL(k_shg_low, dim3(nchunks), dim3(WN_SHG_TPB), 0, st, chunks, count, coef, one, partial,
                           sh.n * sh.n, sp);
    }
    return hipGetLastError();
}
