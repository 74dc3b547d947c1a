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
//   k_shg_mfma  : sum over the chunk's k points and pairs of resonance x coefficient for every
//                 frequency (lane = frequency), in a fixed order, as a real FP64 MFMA GEMM whose A
//                 operand (the resonances) is generated in registers.
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
// per-k blocks of the vertex buffer: der[3a+b] (9), then the regularized dipoles r_eta[a] (3)
#define WN_SHG_DER_BLOCKS 12

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
// der[k][3a+b][n][m] (OpticalPositionDerivatives.jl:13-130) for active pairs, 0 elsewhere.
// Grid (count, 9): block (k, ab).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_der(const double *__restrict__ x, const double *__restrict__ eg,
                                                       double2 *__restrict__ der, WnShape sh, WnShg sp) {
    __shared__ double occ[WN_SHG_NMAX];
    const long long k = blockIdx.x;
    const int ab = blockIdx.y, a = ab / 3, b = ab - (ab / 3) * 3;
    const int n = sh.n, n2 = n * n;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    const int t = threadIdx.x;
    if (ab >= 9) {
        // r_eta[a'] = A + i V reg(E_j - E_i) of every element, read by the pair coefficients' band sums
        double2 *rd = der + ((size_t)k * WN_SHG_DER_BLOCKS + ab) * n2;
        for (int e = t; e < n2; e += WN_SHG_TPB) rd[e] = sv_rdip(s, ab - 9, e / n, e - (e / n) * n);
        return;
    }
    if (t < n) occ[t] = shg_occupation(s.E[t], sp.mu, sp.temperature);
    __syncthreads();
    double2 *out = der + ((size_t)k * WN_SHG_DER_BLOCKS + ab) * n2;
    const bool pos = s.pos != 0;
    for (int e = t; e < n2; e += WN_SHG_TPB) {
        const int nn = e / n, m = e - (e / n) * n;
        if (!shg_active(occ, nn, m)) {
            out[e] = make_double2(0.0, 0.0);
            continue;
        }
        // commutators [A_a, reg_b] and [V_a, reg_b]
        double2 ca = make_double2(0.0, 0.0), ch = make_double2(0.0, 0.0);
        double2 eta_num = make_double2(0.0, 0.0);
        for (int p = 0; p < n; ++p) {
            const double2 rpm = sv_reg(s, b, p, m), rnp = sv_reg(s, b, nn, p);
            if (pos) ca = csub(cadd(ca, cmul(sv_A(s, a, nn, p), rpm)), cmul(rnp, sv_A(s, a, p, m)));
            ch = csub(cadd(ch, cmul(sv_V(s, a, nn, p), rpm)), cmul(rnp, sv_V(s, a, p, m)));
            if (sp.eta_correction && p != nn && p != m && pos) {
                const double wpm = shg_weight(s.E[p] - s.E[m], sp.eta);
                const double wnp = shg_weight(s.E[nn] - s.E[p], sp.eta);
                const double2 Aanp = sv_A(s, a, nn, p), Vbpm = sv_V(s, b, p, m), Abpm = sv_A(s, b, p, m);
                const double2 Vanp = sv_V(s, a, nn, p);
                // V_a,np + i (E_n - E_p) A_a,np
                const double2 van = cadd(Vanp, cmuli(cscale(Aanp, s.E[nn] - s.E[p])));
                eta_num = cadd(eta_num, cscale(csub(cmul(Aanp, Vbpm), cmul(van, Abpm)), wpm));
                const double2 Vbnp = sv_V(s, b, nn, p), Aapm = sv_A(s, a, p, m), Abnp = sv_A(s, b, nn, p);
                const double2 Vapm = sv_V(s, a, p, m);
                const double2 vap = cadd(Vapm, cmuli(cscale(Aapm, s.E[p] - s.E[m])));
                eta_num = csub(eta_num, cscale(csub(cmul(Vbnp, Aapm), cmul(Abnp, vap)), wnp));
            }  // without a position operator every eta-correction product carries an A factor
        }
        const double2 Aan = sv_A(s, a, nn, nn), Aam = sv_A(s, a, m, m);
        const double2 Abn = sv_A(s, b, nn, nn), Abm = sv_A(s, b, m, m);
        const double2 Van = sv_V(s, a, nn, nn), Vam = sv_V(s, a, m, m);
        const double2 Vbn = sv_V(s, b, nn, nn), Vbm = sv_V(s, b, m, m);
        const double2 reg_b = sv_reg(s, b, nn, m);
        const double2 bare_a = sv_bare(s, a, nn, m), bare_b = sv_bare(s, b, nn, m);
        const double2 dAa = csub(Aan, Aam), dAb = csub(Abn, Abm);
        const double2 sum_a = csub(ca, cmul(dAa, reg_b));
        const double2 sum_h = csub(ch, cmul(csub(Van, Vam), reg_b));
        const double2 dAab = pos ? s.dA(3 * a + b)[nn * n + m] : make_double2(0.0, 0.0);
        double2 finite = cadd(dAab, sum_a);
        finite = cadd(finite, cmul(dAa, bare_b));
        finite = cadd(finite, cmul(dAb, bare_a));
        finite = csub(finite, cmuli(cmul(sv_A(s, a, nn, m), dAb)));
        const double2 Wab = s.W(wn_sym(a, b))[nn * n + m];
        double2 inner = cadd(Wab, sum_h);
        inner = cadd(inner, cscale(bare_a, Vbn.x - Vbm.x));
        inner = cadd(inner, cscale(bare_b, Van.x - Vam.x));
        double2 numerator = cadd(cmuli(inner), eta_num);
        const double iv = shg_bare_inverse(s.E[m] - s.E[nn]);
        out[e] = iv == 0.0 ? finite : cadd(finite, cscale(numerator, iv));
    }
}

// ------------------------------------------------------------------------------------------------
// Pair list, coefficient rows and one-band sums of one k point (one workgroup per k).
// pairs[k][p] = (n, m), pdiff[k][p] = E_m - E_n, coef[k][p][col], count[k]; one[k][comp][2] = (bcd, semi).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(WN_SHG_TPB) void k_shg_coef(const double *__restrict__ x, const double *__restrict__ eg,
                                                        const double2 *__restrict__ der, int *__restrict__ count,
                                                        int2 *__restrict__ pairs, double *__restrict__ pdiff,
                                                        double2 *__restrict__ coef, double2 *__restrict__ one,
                                                        WnShape sh, WnShg sp) {
    __shared__ double occ[WN_SHG_NMAX];
    __shared__ int rowcnt[WN_SHG_NMAX];
    __shared__ int rowoff[WN_SHG_NMAX + 1];
    extern __shared__ double lds_one[];  // [n][nc][2] one-band contributions
    const long long k = blockIdx.x;
    const int n = sh.n, n2 = n * n, nc = sp.ncomp;
    const int ncol = 19 * nc;
    const ShgView s = make_view(x, eg, k, sh, sp.eta);
    const double2 *dk = der + (size_t)k * WN_SHG_DER_BLOCKS * n2;
    const double2 *rdk = dk + (size_t)9 * n2;   // r_eta[a][i][j]
    const int t = threadIdx.x;
    if (t < n) occ[t] = shg_occupation(s.E[t], sp.mu, sp.temperature);
    __syncthreads();
    // n-major pair order (SecondHarmonicKernels.jl:140-141)
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
        for (int m = 0; m < n; ++m)
            if (shg_active(occ, t, m)) {
                pairs[pbase + p] = make_int2(t, m);
                pdiff[pbase + p] = s.E[m] - s.E[t];
                ++p;
            }
    }
    __syncthreads();
    const int np = rowoff[n];
    auto D = [&](int aa, int bb, int i, int j) { return dk[(size_t)(3 * aa + bb) * n2 + i * n + j]; };
    for (int item = t; item < np * nc; item += WN_SHG_TPB) {
        const int p = item / nc, ci = item - (item / nc) * nc;
        const int2 pr = pairs[pbase + p];
        const int nn = pr.x, m = pr.y;
        const int cc = sp.comp[ci];
        const int a = cc % 3, b = (cc / 3) % 3, c = cc / 9;
        const double diff = s.E[m] - s.E[nn];
        const double inverse = shg_bare_inverse(diff);
        const double o = occ[nn] - occ[m];
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
        double2 tf = make_double2(0.0, 0.0), to = make_double2(0.0, 0.0);
        for (int q = 0; q < n; ++q) {
            if (q == nn || q == m) continue;
            const double2 rb_mq = rdk[(size_t)b * n2 + m * n + q], rc_qn = rdk[(size_t)c * n2 + q * n + nn];
            const double2 rc_mq = rdk[(size_t)c * n2 + m * n + q], rb_qn = rdk[(size_t)b * n2 + q * n + nn];
            const double pv0 = shg_principal(2.0 * s.E[q] - s.E[nn] - s.E[m], sp.eta);
            tf = cadd(tf, cscale(cmul(dip_a_nm, cadd(cmul(rb_mq, rc_qn), cmul(rc_mq, rb_qn))), pv0));
            const double pv1 = shg_principal(2.0 * s.E[nn] - s.E[m] - s.E[q], sp.eta);
            const double2 ra_mq = rdk[(size_t)a * n2 + m * n + q];
            to = cadd(to, cscale(cmul(ra_mq, cadd(cmul(rb_qn, dip_c_nm), cmul(rc_qn, dip_b_nm))), pv1));
            const double pv2 = shg_principal(2.0 * s.E[m] - s.E[q] - s.E[nn], sp.eta);
            const double2 ra_qn = rdk[(size_t)a * n2 + q * n + nn];
            to = cadd(to, cscale(cmul(ra_qn, cadd(cmul(dip_b_nm, rc_mq), cmul(dip_c_nm, rb_mq))), pv2));
        }
        // f = i o / 2;  f * inv * z = i (o inv / 2) z
        const double h = 0.5 * o;
        const double hi = h * inverse;
        double2 *row = coef + ((size_t)pbase + p) * ncol;
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
    // one-band terms (SecondHarmonicKernels.jl:245-294), T > 1e-7 only
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
// Frequency contraction on FP64 MFMA.  For one chunk and one resonance kind (d1, d2, dn) the
// partial sums form a complex GEMM out[w][c] = sum_p R(dE_p, w) coef[p][c], written as a real GEMM
// with K = 2 x pairs:  A[w][2p] = Re R, A[w][2p+1] = Im R;  B[2p][2c] = Re C, B[2p][2c+1] = Im C,
// B[2p+1][2c] = -Im C, B[2p+1][2c+1] = Re C.  v_mfma_f64_16x16x4f64 lane l holds A[l%16][l/16] (the
// resonance of frequency l%16 and pair l/32 of the step, part (l/16)%2, generated in registers) and
// B[l/16][l%16] (loaded from the coefficient rows).  Workgroup (chunk, 64-frequency tile, kind);
// wave w owns 16 frequencies and every 16-column tile of its kind's block.  Kind 3 is the
// frequency-independent low block: plain sums over the chunk's pairs in pair order.  Coefficient rows
// are staged through LDS in blocks of `pb` pairs (all k points of the chunk in order).
// ------------------------------------------------------------------------------------------------
#define WN_SHG_MAXK 64   // k points per chunk
template <int MAXT, int RTW>
__global__ __launch_bounds__(WN_SHG_TPB, MAXT == 4 ? 3 : 2) void k_shg_mfma(const int4 *__restrict__ chunks, const int *__restrict__ count,
                                                        const double *__restrict__ pdiff,
                                                        const double2 *__restrict__ coef,
                                                        const double2 *__restrict__ one,
                                                        const double *__restrict__ omega, double2 *__restrict__ partial,
                                                        int pcap, int nwt, int pb, WnShg sp) {
    extern __shared__ double lds_b[];          // [pb][2][2 cwid] coefficient block in the real B layout
    __shared__ double lds_diff[128];
    __shared__ int lds_row[128];
    __shared__ int pre[WN_SHG_MAXK + 1];       // prefix of the pair counts of the chunk's k points
    const int nc = sp.ncomp, ncol = 19 * nc, nw = sp.nw;
    const int per_chunk = nwt * 4;
    const int chunk = blockIdx.x / per_chunk;
    const int rem = blockIdx.x - chunk * per_chunk;
    const int wt = rem >> 2, kind = rem & 3;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int4 ch = chunks[chunk];
    double2 *o = partial + (size_t)chunk * ((size_t)ncol * nw + nc);
    if (kind == 3) {
        // low block (R = 1) and the one-band sums, once per chunk: thread t sums the pairs t, t + 256,
        // ... of every column, then the 256 partial sums are combined in a fixed tree order
        if (wt != 0) return;
        __shared__ double2 red[WN_SHG_TPB];
        if (t == 0) {
            pre[0] = 0;
            for (int i = 0; i < ch.y; ++i) pre[i + 1] = pre[i] + count[ch.x + i];
        }
        __syncthreads();
        const int tot = pre[ch.y];
        for (int c = 0; c < 3 * nc; ++c) {
            double2 acc = make_double2(0.0, 0.0);
            int ki = 0;
            for (int f = t; f < tot; f += WN_SHG_TPB) {
                while (pre[ki + 1] <= f) ++ki;
                acc = cadd(acc, coef[((size_t)(ch.x + ki) * pcap + (f - pre[ki])) * ncol + 16 * nc + c]);
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
        return;
    }
    const int cbeg = kind == 0 ? 0 : (kind == 1 ? 8 * nc : 14 * nc);
    const int cwid = kind == 0 ? 8 * nc : (kind == 1 ? 6 * nc : 2 * nc);   // complex columns
    const int ntile = (2 * cwid + 15) / 16;
    const int lr = lane & 15, lk = lane >> 4;           // A row (frequency) / K row of this lane
    // row tiles (16 frequencies each) of this wave: wt * 4 RTW + wave RTW + r
    const int rt0 = (wt * 4 + wave) * RTW;
    double om[RTW];
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
        const int w_a = (rt0 + r) * 16 + lr;            // frequency of this lane's A element in tile r
        om[r] = w_a < nw ? omega[w_a] : 0.0;
    }
    const int pair_sel = lk >> 1, part = lk & 1;        // pair of the step and re / im row
    if (t == 0) {
        pre[0] = 0;
        for (int i = 0; i < ch.y; ++i) pre[i + 1] = pre[i] + count[ch.x + i];
    }
    __syncthreads();
    const int total = pre[ch.y];
    f64x4 acc[RTW][MAXT];
#pragma unroll
    for (int r = 0; r < RTW; ++r)
#pragma unroll
        for (int j = 0; j < MAXT; ++j) acc[r][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
    const double cut = 27.3 * sp.width;                 // exp(-(x/w)^2) is 0 in FP64 beyond this
    const double w2 = sp.width * sp.width, iw = 1.0 / sp.width;
    const double gnorm = sqrt(M_PI) * iw;               // pi / (sqrt(pi) w)
    for (int f0 = 0; f0 < total; f0 += pb) {
        const int nb = min(pb, total - f0);
        // cooperative load of the block's coefficient rows and energy differences (pairs in chunk order);
        // the row of every pair of the block is located once
        for (int e = t; e < nb; e += WN_SHG_TPB) {
            const int f = f0 + e;
            int ki = 0;
            while (pre[ki + 1] <= f) ++ki;
            const int row = (ch.x + ki) * pcap + (f - pre[ki]);
            lds_row[e] = row;
            lds_diff[e] = pdiff[row];
        }
        __syncthreads();
        // pair fp, coefficient c = (cr, ci): B rows 2fp = (cr, ci), 2fp+1 = (-ci, cr)
        const int bw = 2 * cwid;
        for (int e = t; e < nb * cwid; e += WN_SHG_TPB) {
            const int fp = e / cwid, c = e - (e / cwid) * cwid;
            const double2 cv = coef[(size_t)lds_row[fp] * ncol + cbeg + c];
            double *b0 = lds_b + (size_t)(2 * fp) * bw + 2 * c;
            b0[0] = cv.x;
            b0[1] = cv.y;
            b0[bw] = -cv.y;
            b0[bw + 1] = cv.x;
        }
        __syncthreads();
#pragma unroll 2
        for (int s0 = 0; s0 < nb; s0 += 2) {
            const int pl = s0 + pair_sel;
            const bool valid = pl < nb;
            // A: resonances R(x) = PV(x) + i pi delta(x) (SecondHarmonicKernels.jl:86-98) of (frequency of row
            // tile r, pair pl); part 0 of the lane carries Re R, part 1 Im R
            const double diff = valid ? lds_diff[pl] : 0.0;
            double a[RTW];
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
                const double xx = kind == 0 ? diff - om[r] : (kind == 1 ? diff - 2.0 * om[r] : -diff - om[r]);
                double v;
                if (part == 0) {
                    v = xx * wn_rcp(xx * xx + w2);                      // PV: x / (x^2 + w^2)
                } else if (sp.gaussian) {
                    const double y = xx * iw;                           // pi exp(-y^2) / (sqrt(pi) w)
                    v = fabs(xx) < cut ? gnorm * exp(-y * y) : 0.0;
                } else {
                    const double y = xx * iw;                           // pi / (pi w (1 + y^2))
                    v = iw * wn_rcp(1.0 + y * y);
                }
                a[r] = valid ? v : 0.0;
            }
            const double *brow = lds_b + (size_t)(2 * (valid ? pl : 0) + part) * bw;
#pragma unroll
            for (int j = 0; j < MAXT; ++j) {
                if (j < ntile) {
                    const int cr = j * 16 + lr;                  // real column
                    const double b = (valid && cr < bw) ? brow[cr] : 0.0;
#pragma unroll
                    for (int r = 0; r < RTW; ++r)
                        acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b, acc[r][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D reg q -> row (frequency) lk + 4q, column lr of each 16-column tile
#pragma unroll
    for (int r = 0; r < RTW; ++r)
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        if (j < ntile) {
            const int cr = j * 16 + lr;
            if (cr < 2 * cwid) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int w = (rt0 + r) * 16 + lk + 4 * q;
                    if (w < nw) {
                        double *dst = (double *)(o + (size_t)(cbeg + (cr >> 1)) * nw + w);
                        dst[cr & 1] = acc[r][j][q];
                    }
                }
            }
        }
    }
}

hipError_t wn_launch_shg(const double *x, const double *e, const double *omega, const int4 *chunks, int nchunks,
                         int count_k, double2 *der, int *count, int2 *pairs, double *pdiff, double2 *coef,
                         double2 *one, double2 *partial, const WnShape &sh, const WnShg &sp, hipStream_t st) {
    if (sh.n > WN_SHG_NMAX || sp.ncomp < 1 || sp.ncomp > 18) return hipErrorInvalidValue;
    if (count_k > 0) {
        hipLaunchKernelGGL(k_shg_der, dim3(count_k, WN_SHG_DER_BLOCKS), dim3(WN_SHG_TPB), 0, st, x, e, der, sh, sp);
        hipLaunchKernelGGL(k_shg_coef, dim3(count_k), dim3(WN_SHG_TPB), (size_t)sh.n * sp.ncomp * 2 * sizeof(double),
                           st, x, e, der, count, pairs, pdiff, coef, one, sh, sp);
    }
    if (nchunks > 0) {
        // RTW row tiles of 16 frequencies per wave: a workgroup covers 64 RTW frequencies
        const int rtw = sp.ncomp <= 4 ? 3 : (sp.ncomp <= 8 ? 2 : 1);
        const int nwt = (sp.nw + 64 * rtw - 1) / (64 * rtw);
        // pairs per LDS block: the widest block (8 nc complex columns) times pb within 32 KB, even, <= 128
        // pairs per LDS block: the widest block (8 nc complex columns, 4 doubles each in the B layout) within
        // 32 KB, even, <= 128
        const int pb = std::max(2, std::min(128, (1024 / (8 * sp.ncomp)) & ~1));
        const size_t lds = (size_t)pb * 8 * sp.ncomp * 4 * sizeof(double);
        const dim3 grid(nchunks * nwt * 4);
#define WN_SHG_LAUNCH(T, R)                                                                                         \
    hipLaunchKernelGGL((k_shg_mfma<T, R>), grid, dim3(WN_SHG_TPB), lds, st, chunks, count, pdiff, coef, one, omega, \
                       partial, sh.n * sh.n, nwt, pb, sp)
        if (sp.ncomp <= 4)
            WN_SHG_LAUNCH(4, 3);
        else if (sp.ncomp <= 8)
            WN_SHG_LAUNCH(8, 2);
        else
            WN_SHG_LAUNCH(18, 1);
#undef WN_SHG_LAUNCH
    }
    return hipGetLastError();
}
