// Batched Hermitian Jacobi eigensolver with register-resident eigenvectors (n_w <= 32).
//
// Same mathematics and sweep order as the LDS kernel k_eig_f64 (wnlqg_kernels.hip): the upper triangle
// of channel 0 is the Hermitian matrix (DenseSpectra.hermitian_spectrum, reference
// src/Core/ESKernel/DenseSpectra.jl:14-16, uplo 'U'), parallel cyclic Jacobi over a round-robin
// tournament (element 0 fixed, the others moving one position per round), the 2x2 block form of a round
// on the packed upper triangle, the same stopping rule, ascending sort and canonical column phase
// (canonicalize_eigenvectors!, HamiltonianMatrixElements.jl:6-25).  What differs is where the work lives:
//
//   * A workgroup of two waves serves KPW = 64 / N k points (N = 8, 16, 32: the padded size).
//   * Wave 0 owns H: both packed triangles in LDS, N lanes per k point.  It computes the rotation
//     parameters of a round (lanes 0 .. N/2-1 of each k point's lane group), then the off-diagonal block
//     updates of that round, then the parameters of the next round: every read of H follows the writes
//     it depends on in the same wave's program order, so H needs no workgroup barrier.
//   * Wave 1 owns V: lane i of a k point's group holds row i of V (N complex values) in registers.
//     V <- V J combines two columns of the same row, so a row never leaves its lane; the round loop is
//     unrolled, which makes every round's column pairs compile-time register indices.
//   * One barrier per round hands the round's rotation parameters (double-buffered in LDS) from wave 0
//     to wave 1.  A k point that has converged gets identity rotations, which leave H and V unchanged
//     exactly, so its result is what a solve of that point alone gives.
//
// V's LDS traffic (every rotated column pair read and written each round) is gone; what remains in
// LDS is the packed H (read and written once per block per round) and the parameter broadcast.

#include <algorithm>

#include "wnlqg_common.h"
#include "wnlqg_kernels.h"
#include "wnlqg_eigd_lds32.h"

namespace {

// Orders this wave's LDS accesses in program order for the compiler (the LDS executes one wave's
// instructions in issue order, so cross-lane reads after writes of the same wave need no wait).
__device__ __forceinline__ void wn_wave_order() { __asm__ volatile("" ::: "memory"); }

__device__ __forceinline__ void wn_wg_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Round-robin tournament of N players in N - 1 rounds (circle method): slot t of round r holds
//   left a = 0 (t = 0) or (t - 1 + r) mod (N - 1) + 1,   right b = (N - 2 - t + r) mod (N - 1) + 1,
// the LDS kernel's pairs.  The rotation of a slot acts on (left, right) in that order.
__host__ __device__ constexpr int wn_rr_a(int N, int t, int r) {
    return t == 0 ? 0 : (t - 1 + r) % (N - 1) + 1;
}
__host__ __device__ constexpr int wn_rr_b(int N, int t, int r) { return (N - 2 - t + r) % (N - 1) + 1; }
// (left, right) of slot t in round r at run time: one conditional subtraction per residue
__device__ __forceinline__ int2 wn_rr_rt(int N, int t, int r) {
    int a = t - 1 + r;
    a = a >= N - 1 ? a - (N - 1) : a;
    int b = N - 2 - t + r;
    b = b >= N - 1 ? b - (N - 1) : b;
    return make_int2(t == 0 ? 0 : a + 1, b + 1);
}

// A row of V held as register tuples of 8 complex values (16 doubles): element updates are writes into the
// same tuple, so the row keeps its registers from one sweep to the next.
typedef double f64x16 __attribute__((ext_vector_type(16)));
template <int N>
struct WnRow {
    f64x16 t[(N + 7) / 8];
    template <int C>
    __device__ __forceinline__ double2 get() const {
        return make_double2(t[C / 8][2 * (C % 8)], t[C / 8][2 * (C % 8) + 1]);
    }
    template <int C>
    __device__ __forceinline__ void set(double2 v) {
        t[C / 8][2 * (C % 8)] = v.x;
        t[C / 8][2 * (C % 8) + 1] = v.y;
    }
};

// V <- V J for one row: columns (p, q) with R = (c, s, cos phi, sin phi)
template <int N, int P, int Q>
__device__ __forceinline__ void wn_rotate_row(WnRow<N> &V, const double4 R) {
    const double2 xp = V.template get<P>(), yq = V.template get<Q>();
    const double2 ey = make_double2(R.z * yq.x + R.w * yq.y, R.z * yq.y - R.w * yq.x);
    V.template set<P>(make_double2(R.x * xp.x - R.y * ey.x, R.x * xp.y - R.y * ey.y));
    V.template set<Q>(make_double2(R.y * xp.x + R.x * ey.x, R.y * xp.y + R.x * ey.y));
}

// One unrolled round of the V wave: every slot's rotation applied to this lane's row.
// The parameters are loaded four slots at a time (a compiler fence between groups), so that the row and
// the parameters in flight stay within the register budget of two waves per SIMD.
template <int N, int R, int T, int G = 4>
struct WnVSlots {
    __device__ __forceinline__ static void run(WnRow<N> &V, const double4 *rot) {
        // parameters loaded G slots at a time (a compiler fence between groups)
        if constexpr (T % G == 0) wn_wave_order();
        wn_rotate_row<N, wn_rr_a(N, T, R), wn_rr_b(N, T, R)>(V, rot[T]);
        WnVSlots<N, R, T + 1, G>::run(V, rot);
    }
};
template <int N, int R, int G>
struct WnVSlots<N, R, N / 2, G> {
    __device__ __forceinline__ static void run(WnRow<N> &, const double4 *) {}
};

// The V wave's rounds of one sweep; the round's parameters are in rot[(cb + R) & 1].  The barrier of round 0
// also publishes the go flag, read there once (the H wave rewrites it for the next sweep while this wave may
// still be in its last round): without another sweep the V wave stops there and run returns false.
template <int N, int R>
struct WnVRounds {
    __device__ __forceinline__ static bool run(WnRow<N> &V, const double4 (*rot)[64 / N][N / 2], int j, int cb,
                                               const volatile int &go) {
        wn_wg_barrier();
        if (R == 0 && !go) return false;
        WnVSlots<N, R, 0>::run(V, rot[(cb + R) & 1][j]);
        return WnVRounds<N, R + 1>::run(V, rot, j, cb, go);
    }
};
template <int N>
struct WnVRounds<N, N - 1> {
    __device__ __forceinline__ static bool run(WnRow<N> &, const double4 (*)[64 / N][N / 2], int, int,
                                               const volatile int &) {
        return true;
    }
};

template <int N, int C>
struct WnRowIo {
    __device__ __forceinline__ static void load(WnRow<N> &V, const double2 *__restrict__ ua, int n, int l) {
        V.template set<C>((ua && l < n && C < n) ? ua[l * n + C] : make_double2(l == C ? 1.0 : 0.0, 0.0));
        WnRowIo<N, C + 1>::load(V, ua, n, l);
    }
    // row l as it is (columns in band order)
    __device__ __forceinline__ static void store_raw(const WnRow<N> &V, double2 *__restrict__ uk, int n, int l) {
        if (C < n) uk[l * n + C] = V.template get<C>();
        WnRowIo<N, C + 1>::store_raw(V, uk, n, l);
    }
    // canonical phase of column C (first largest |V[i][C]| over the rows i < n made real and >= 0), written
    // to column rank[C] of the output
    __device__ __forceinline__ static void store(const WnRow<N> &V, double2 *__restrict__ uk, const int *rank, int n,
                                                 int l, int base) {
        if (C < n) {   // uniform over the wave
            const double2 v = V.template get<C>();
            double m = l < n ? hypot(v.x, v.y) : -1.0;
            int idx = l;
#pragma unroll
            for (int o = N / 2; o > 0; o >>= 1) {
                const double om = __shfl_xor(m, o);
                const int oi = __shfl_xor(idx, o);
                if (om > m || (om == m && oi < idx)) {
                    m = om;
                    idx = oi;
                }
            }
            const double2 pv = make_double2(__shfl(v.x, base + idx), __shfl(v.y, base + idx));
            const double2 ph = m > 0.0 ? make_double2(pv.x / m, -pv.y / m) : make_double2(1.0, 0.0);
            if (l < n) uk[l * n + rank[C]] = cmul(v, ph);
        }
        WnRowIo<N, C + 1>::store(V, uk, rank, n, l, base);
    }
};
template <int N>
struct WnRowIo<N, N> {
    __device__ __forceinline__ static void load(WnRow<N> &, const double2 *__restrict__, int, int) {}
    __device__ __forceinline__ static void store(const WnRow<N> &, double2 *__restrict__, const int *, int, int, int) {}
    __device__ __forceinline__ static void store_raw(const WnRow<N> &, double2 *__restrict__, int, int) {}
};

}  // namespace

// Slot-ordered storage of H.  Round r of the tournament pairs the bands (a_t(r), b_t(r)) in slot t; H is
// kept with the bands of slot t at positions (2t, 2t+1) (left, right), so that slot t's 2x2 diagonal block
// and every off-diagonal block (t1, t2) sit at the same addresses in every round.  Between rounds the
// circle method moves every band one place along the cycle R_0 -> R_1 -> ... -> R_{H-1} -> L_{H-1} -> ...
// -> L_1 -> R_0 (L_t = 2t, R_t = 2t + 1, band 0 fixed at L_0): the same position map sigma every round.
// A round reads its blocks from one buffer and writes each updated element to (sigma(x), sigma(y)) of the
// other buffer (conjugated when that lands below the diagonal), so every address and conjugation of a
// lane is fixed for the whole solve.  Storage: upper-triangle 2x2 blocks (t1 <= t2) row-major, element
// (u, v) of block b at [2u + v][b] (each element plane contiguous over blocks).  After a whole sweep
// (N - 1 rounds) every band is back at its round-0 position pos0(b).
template <int N>
struct WnSlots {
    static constexpr int HALF = N / 2;
    static constexpr int NB = HALF * (HALF + 1) / 2;
    __host__ __device__ static constexpr int blk(int t1, int t2) { return t1 * HALF - t1 * (t1 - 1) / 2 + (t2 - t1); }
    __host__ __device__ static constexpr int sigma(int x) {
        return x == 0 ? 0
                      : ((x & 1) ? ((x >> 1) < HALF - 1 ? x + 2 : x - 1) : ((x >> 1) == 1 ? 1 : x - 2));
    }
    __host__ __device__ static constexpr int pos0(int b) { return b < HALF ? 2 * b : 2 * (N - 1 - b) + 1; }
    // storage index of element (x, y), x <= y
    __host__ __device__ static constexpr int at(int x, int y) { return (2 * (x & 1) + (y & 1)) * NB + blk(x >> 1, y >> 1); }
    // destination of the element now at (x, y), x < y: index, with bit 31 set when stored conjugated
    __host__ __device__ static constexpr int dest(int x, int y) {
        return sigma(x) < sigma(y) ? at(sigma(x), sigma(y)) : (at(sigma(y), sigma(x)) | (int)0x80000000);
    }
};

template <int N>
struct WnEigShared {
    double2 Hs[2][64 / N][4 * WnSlots<N>::NB];   // slot-ordered H, two buffers
    double4 rot[2][64 / N][N / 2];               // (c, s, cos phi, sin phi) of the round, two buffers
    int rank[64 / N][N];                          // position of band c in the ascending spectrum
    int go_flag;                                  // another sweep follows
};

// Rotation parameters of one slot (branch-free, so that the compiler can interleave the dependent chain
// with independent block updates): H's 2x2 diagonal block [[a, h], [h*, d]] of the (left, right) pair.
// rotate = false gives the identity (1, 0, 1, 0) and leaves the block as it is.
struct WnRotParams {
    double4 R;       // (c, s, cos phi, sin phi)
    double an, dn;   // new diagonal entries
    double2 h;       // new (left, right) entry: 0 after a rotation
};
__device__ __forceinline__ WnRotParams wn_rot_params(double a, double d, double2 h, bool rotate) {
    // smaller root t of t^2 + 2 tau t - 1 = 0, tau = (d - a)/(2b): with root = sqrt((d-a)^2 + 4b^2) and
    // u = |d-a| + root: c = u / sqrt(2 root u), s = sign(d-a) 2b / sqrt(2 root u), t b = sign(d-a) 2b^2 / u
    const double b2 = h.x * h.x + h.y * h.y;
    rotate = rotate && b2 > 0.0;
    const double b2s = rotate ? b2 : 1.0;
    const double ib = wn_rsqrt(b2s);
    const double diff = d - a;
    const double rr = diff * diff + 4.0 * b2s;
    const double root = rr * wn_rsqrt(rr);
    const double u = fabs(diff) + root;
    const double iw = wn_rsqrt(2.0 * root * u);
    const double sg = diff >= 0.0 ? 1.0 : -1.0;
    const double tb = sg * 2.0 * b2s * wn_rcp(u);
    WnRotParams o;
    o.R = rotate ? make_double4(u * iw, sg * 2.0 * (b2s * ib) * iw, h.x * ib, h.y * ib)
                 : make_double4(1.0, 0.0, 1.0, 0.0);
    o.an = rotate ? a - tb : a;
    o.dn = rotate ? d + tb : d;
    o.h = rotate ? make_double2(0.0, 0.0) : h;
    return o;
}

// One off-diagonal block B (t1 < t2) of round buffer Hr: B <- J1^+ B J2, its four elements written to the
// next round's positions dst[e] (bit 31: stored conjugated) of Hw.
__device__ __forceinline__ void wn_block_update(const double2 *Hr, double2 *Hw, int src, const int (&dst)[4], int NB,
                                                const double4 R1, const double4 R2) {
    double2 B[2][2];
#pragma unroll
    for (int e = 0; e < 4; ++e) B[e >> 1][e & 1] = Hr[e * NB + src];
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
    for (int e = 0; e < 4; ++e) {
        const int dd = dst[e];
        const double2 val = B[e >> 1][e & 1];
        Hw[dd & 0x7fffffff] = make_double2(
            val.x, __longlong_as_double(__double_as_longlong(val.y) ^ ((unsigned long long)(dd < 0) << 63)));
    }
}

// Wave 0: H, the stopping test, the rotation parameters, the eigenvalues.
//
// Look-ahead: the diagonal blocks of round r + 1 come from round r's diagonal entries and from one
// off-diagonal block per slot, the "critical" block whose (L_{t+1}, R_{t-1}) element becomes slot t's
// (left, right) entry: (0, 1) for slot 0, (t-1, t+1) for 1 <= t <= H-2, (H-2, H-1) for slot H-1.  A round
// updates the critical blocks first, then computes the next round's parameters (published before the
// next barrier), then the other blocks; the V wave can start round r + 1 while those finish.  The
// parameter step writes only diagonal entries and critical-block positions of the buffer being read,
// which the other blocks do not read.  Barriers: N - 1 per sweep (the first also carries the go flag),
// one at the end, as in wn_eig_v_role.
template <int N>
__device__ __forceinline__ void wn_eig_h_role(WnEigShared<N> &sm, const double *__restrict__ x,
                                              double *__restrict__ eout, int *__restrict__ sweeps_out,
                                              unsigned long long *__restrict__ sweep_total, int n, int ldo,
                                              double tol, int k, bool active, int j, int l, int lane, int bucket,
                                              int *__restrict__ status, int max_sweeps,
                                              const double2 *__restrict__ hsrc = nullptr, int sweep0 = 0,
                                              double prev0 = 0.0) {
    using S = WnSlots<N>;
    constexpr int HALF = N / 2;
    constexpr int NB = S::NB;
    constexpr int NOFF = HALF * (HALF - 1) / 2;
    constexpr int NREST = NOFF - HALF;                  // off-diagonal blocks that are not critical
    // block slots of a lane: slot 0 is slot cs's critical block on lanes l < HALF and one of the first
    // HALF non-critical blocks on the others; the remaining non-critical blocks fill BPL more slots
    constexpr int BPL = (NREST - HALF + N - 1) / N;
    // H(k) upper triangle in band order -> slot order of round 0; or, continuing a solve, the slot-ordered H
    // saved at the start of sweep sweep0 (every band at pos0) into the buffer that sweep reads
    if (hsrc) {
        for (int e = l; e < 4 * WnSlots<N>::NB; e += N)
            sm.Hs[sweep0 & 1][j][e] = active ? hsrc[e] : make_double2(0.0, 0.0);
    }
    const double2 *xk = active && !hsrc ? (const double2 *)(x + (size_t)k * ldo) : nullptr;
    for (int e = l; e < N * N && !hsrc; e += N) {
        const int i = e / N, jj = e - (e / N) * N;
        if (i <= jj) {
            double2 hv = make_double2(0.0, 0.0);
            if (active && jj < n) {
                hv = xk[i * n + jj];
                if (i == jj) hv.y = 0.0;
            }
            const int pi = S::pos0(i), pj = S::pos0(jj);
            if (pi <= pj)
                sm.Hs[0][j][S::at(pi, pj)] = hv;
            else
                sm.Hs[0][j][S::at(pj, pi)] = make_double2(hv.x, -hv.y);
        }
    }
    // fixed addresses of this lane.  Critical block of next-round slot cs = l mod HALF (lanes l < HALF);
    // the other blocks in row-major order over the strict upper triangle of blocks, critical ones skipped.
    auto critical = [](int t1, int t2) {
        return (t1 == 0 && t2 == 1) || (t2 == t1 + 2) || (t1 == HALF - 2 && t2 == HALF - 1);
    };
    const int cs = l & (HALF - 1);
    // the want-th non-critical block (t1, t2), or (-1, -1)
    auto rest_block = [&](int want, int &t1, int &t2) {
        t1 = -1;
        t2 = -1;
        if (want < 0 || want >= NREST) return;
        int seen = 0;
        for (int a = 0; a < HALF - 1 && t1 < 0; ++a)
            for (int b = a + 1; b < HALF; ++b) {
                if (critical(a, b)) continue;
                if (seen == want) {
                    t1 = a;
                    t2 = b;
                    break;
                }
                ++seen;
            }
    };
    int ct1, ct2;
    if (l < HALF) {
        ct1 = cs == 0 ? 0 : (cs == HALF - 1 ? HALF - 2 : cs - 1);
        ct2 = cs == 0 ? 1 : (cs == HALF - 1 ? HALF - 1 : cs + 1);
    } else {
        rest_block(l - HALF, ct1, ct2);
    }
    const int csrc = ct1 >= 0 ? S::blk(ct1, ct2) : 0;
    int cdst[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) cdst[e] = ct1 >= 0 ? S::dest(2 * ct1 + (e >> 1), 2 * ct2 + (e & 1)) : 0;
    int bt1[BPL > 0 ? BPL : 1], bt2[BPL > 0 ? BPL : 1], bsrc[BPL > 0 ? BPL : 1], bdst[BPL > 0 ? BPL : 1][4];
#pragma unroll
    for (int qb = 0; qb < BPL; ++qb) {
        const int want = HALF + l + qb * N;   // index among the non-critical blocks
        int t1 = -1, t2 = -1;
        rest_block(want, t1, t2);
        bt1[qb] = t1;
        bt2[qb] = t2;
        bsrc[qb] = t1 >= 0 ? S::blk(t1, t2) : 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) bdst[qb][e] = t1 >= 0 ? S::dest(2 * t1 + (e >> 1), 2 * t2 + (e & 1)) : 0;
    }
    // slot cs's diagonal block and where its entries go
    const int dblk = S::blk(cs, cs);
    const int dst_a = S::at(S::sigma(2 * cs), S::sigma(2 * cs));
    const int dst_d = S::at(S::sigma(2 * cs + 1), S::sigma(2 * cs + 1));
    const int dst_h = S::dest(2 * cs, 2 * cs + 1) & 0x7fffffff;
    const bool writer = l < HALF;

    // parameters of round r (buffer Hr) for slot cs, published in rot[g & 1], the diagonal block written to Hw
    auto params = [&](const double2 *Hr, double2 *Hw, int r, int g, bool frz) {
        const int2 pq = wn_rr_rt(N, cs, r);
        const WnRotParams o = wn_rot_params(Hr[dblk].x, Hr[3 * NB + dblk].x, Hr[NB + dblk],
                                            !frz && pq.x < n && pq.y < n);
        if (writer) {
            if (!frz) {
                Hw[dst_a] = make_double2(o.an, 0.0);
                Hw[dst_d] = make_double2(o.dn, 0.0);
                Hw[dst_h] = o.h;
            }
            sm.rot[g & 1][j][cs] = o.R;
        }
    };

    int sweep = sweep0, done_sweep = 0;
    bool frozen = !active, converged = !active;
    double prev_off = prev0;
    for (;;) {
        const int cb = sweep & 1;   // buffer holding H at the start of this sweep (N - 1 rounds per sweep, odd)
        const int g0 = sweep * (N - 1);
        // stopping test of every k point (its N lanes): ||offdiag||_F against tol ||H||_F, or the rounding
        // floor (a sweep no longer reduces the off-diagonal mass once it is ~1e-20 of the total)
        double off = 0.0, dia = 0.0;
        if (!frozen) {
            for (int b = l; b < NB; b += N) {
                int t1 = 0, rem = b;
                while (rem >= HALF - t1) {
                    rem -= HALF - t1;
                    ++t1;
                }
                const double2 e0 = sm.Hs[cb][j][b], e1 = sm.Hs[cb][j][NB + b], e3 = sm.Hs[cb][j][3 * NB + b];
                const double m0 = e0.x * e0.x + e0.y * e0.y, m1 = e1.x * e1.x + e1.y * e1.y;
                const double m3 = e3.x * e3.x + e3.y * e3.y;
                if (rem == 0) {   // diagonal block: (0,0), (1,1) on the diagonal, (0,1) twice off it
                    dia += m0 + m3;
                    off += 2.0 * m1;
                } else {
                    const double2 e2 = sm.Hs[cb][j][2 * NB + b];
                    off += 2.0 * (m0 + m1 + m3 + e2.x * e2.x + e2.y * e2.y);
                }
            }
        }
#pragma unroll
        for (int o = N / 2; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            dia += __shfl_xor(dia, o);
        }
        if (!frozen && (off <= tol * tol * (dia + off) ||
                        (sweep > 0 && off > 0.25 * prev_off && off <= 1e-20 * (dia + off)))) {
            frozen = true;
            converged = true;
            done_sweep = sweep;
        }
        prev_off = off;
        const bool more = __ballot(!frozen) != 0ull && sweep < max_sweeps;
        if (lane == 0) sm.go_flag = more ? 1 : 0;
        if (more) {
            wn_wave_order();
            params(sm.Hs[cb][j], sm.Hs[cb ^ 1][j], 0, g0, frozen);
        }
        for (int r = 0; r < N - 1; ++r) {
            wn_wg_barrier();
            if (r == 0 && !more) break;
            const int rb = (cb + r) & 1;   // buffer read in this round; the other one is written
            const double2 *Hr = sm.Hs[rb][j];
            double2 *Hw = sm.Hs[rb ^ 1][j];
            const double4 *R = sm.rot[(g0 + r) & 1][j];
            if (!frozen) {
                if (ct1 >= 0) wn_block_update(Hr, Hw, csrc, cdst, NB, R[ct1], R[ct2]);
                wn_wave_order();
            }
            if (r + 1 < N - 1) params(Hw, sm.Hs[rb][j], r + 1, g0 + r + 1, frozen);
            if (!frozen) {
#pragma unroll
                for (int qb = 0; qb < BPL; ++qb)
                    if (bt1[qb] >= 0) wn_block_update(Hr, Hw, bsrc[qb], bdst[qb], NB, R[bt1[qb]], R[bt2[qb]]);
            }
            wn_wave_order();
        }
        if (!more) break;
        ++sweep;
    }
    // eigenvalues ascending (stable by index) and their ranks.  A point's H stopped moving at the sweep it
    // converged (or at the cap), with every band at pos0.
    const int fb = (converged ? done_sweep : sweep) & 1;
    if (active && l < n) {
        auto diag = [&](int b) {
            const int p = S::pos0(b);
            return sm.Hs[fb][j][S::at(p, p)].x;
        };
        const double et = diag(l);
        int rk = 0;
        for (int jj = 0; jj < n; ++jj) {
            const double ej = diag(jj);
            rk += (ej < et) || (ej == et && jj < l);
        }
        sm.rank[j][l] = rk;
        eout[(size_t)k * n + rk] = et;
    }
    if (active && l == 0) {
        const int sw = converged ? done_sweep : sweep;
        if (sweeps_out) atomicMax(sweeps_out + bucket, sw);
        if (sweep_total) atomicAdd(sweep_total + bucket / 2, (unsigned long long)sw);
        if (!converged && status) atomicOr(status + 1, 1);
    }
    wn_wg_barrier();
}

// Wave 1: the rows of V in registers; applies every round's rotations, then writes the canonical
// eigenvectors in ascending order.
template <int N>
__device__ __forceinline__ void wn_eig_v_role(WnEigShared<N> &sm, double2 *__restrict__ uout,
                                              const double2 *__restrict__ ua, int n, int k, bool active, int j,
                                              int l, int sweep0 = 0) {
    WnRow<N> V;
    WnRowIo<N, 0>::load(V, ua, n, l);
    for (int sweep = sweep0;; ++sweep)
        if (!WnVRounds<N, 0>::run(V, sm.rot, j, sweep & 1, sm.go_flag)) break;
    wn_wg_barrier();
    if (!active) return;
    WnRowIo<N, 0>::store(V, uout + (size_t)k * n * n, sm.rank[j], n, l, j * N);
}

// pass < 0: every point from the identity; pass 0 / 1: group anchors, then the other points of each group
// from their anchor's eigenvectors (wn_warm_k); pass 2: every point from uprev[k].
template <int N>
__global__ __launch_bounds__(128, 2) void k_eig_rv(const double *__restrict__ x, double *__restrict__ eout,
                                                   double2 *__restrict__ uout, int *__restrict__ sweeps_out,
                                                   unsigned long long *__restrict__ sweep_total, int n, int ldo,
                                                   double tol, int pass, int count, int items,
                                                   const double2 *__restrict__ uprev, int ctr_stride,
                                                   int *__restrict__ status, int max_sweeps) {
    constexpr int KPW = 64 / N;   // k points per workgroup
    __shared__ WnEigShared<N> sm;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = lane / N, l = lane - (lane / N) * N;   // k slot of the workgroup, lane within its group
    const int item = blockIdx.x * KPW + j;
    int anchor = -1, k = -1;
    if (item < items) {
        if (pass < 0 || pass == 2) {
            k = item;
        } else {
            // wn_warm_k (wnlqg_kernels.hip): anchors first, then the other points of each group of 4
            constexpr int G = 4, A = 1;
            if (pass == 0) {
                const int base = item * G;
                k = base + min(A, count - 1 - base);
            } else {
                const int g = item / (G - 1), jj = item - g * (G - 1);
                const int base = g * G;
                const int al = min(A, count - 1 - base);
                anchor = base + al;
                k = base + (jj < al ? jj : jj + 1);
            }
        }
    }
    const bool active = k >= 0 && k < count;
    if (wave == 0) {
        const int bucket = ctr_stride * (int)(item % WN_CTR_BUCKETS);
        wn_eig_h_role<N>(sm, x, eout, sweeps_out, sweep_total, n, ldo, tol, k, active, j, l, lane, bucket, status,
                         max_sweeps);
    } else {
        const double2 *ua = !active ? nullptr
                                    : (pass == 2 ? uprev + (size_t)k * n * n
                                                 : (anchor >= 0 ? uout + (size_t)anchor * n * n : nullptr));
        wn_eig_v_role<N>(sm, uout, ua, n, k, active, j, l);
    }
}

// ------------------------------------------------------------------------------------------------
// Deferred-eigenvector Jacobi for n_w <= 32 (the default for these sizes).
//
// The register-V kernel above keeps each workgroup's H wave and V wave in lock step (one barrier per round)
// and needs the V rows' 128 VGPRs in every wave of the kernel and two H buffers in LDS, so a CU holds only 8
// points at N = 32 and the round latency is exposed.  Here the same rounds run in two kernels:
//
//   * k_eigd_sweeps: H alone, 64 / N points per wave (N lanes each), no barrier.  H is
//     kept slot-ordered in ONE LDS buffer: every lane reads all the elements its blocks and its diagonal
//     block need before any lane writes (the LDS executes one wave's accesses in issue order), so the moved
//     elements can be written over the old ones.  Each round's rotation parameters go to the record buffer.
//     Same parameters, same block updates, same stopping rule as k_eig_rv, so H evolves identically.
//   * k_eigd_vapply: the recorded rounds applied to the rows of V in registers (one lane per row, the round
//     loop unrolled for compile-time column indices, exactly k_eig_rv's V wave), the records streamed through
//     a four-deep register pipeline and an LDS stage; then the canonical sorted eigenvectors.
//
// At most WN_EIGD_SWEEPS sweeps (WnShape::rec_sweeps) are recorded per point.  A point that needs more (rare: warm starts take 4 to
// 5 sweeps, cold starts 7 to 8 at n = 32) saves its slot-ordered H, and after the V kernel has applied the
// recorded sweeps, k_eig_rv_cont finishes it with the register-V kernel's roles from that state, so every
// point follows the same sequence of rotations whatever path finishes it.
// ------------------------------------------------------------------------------------------------
#define WN_EIGD_SWEEPS 8
#ifndef WN_EIGD_VG
#define WN_EIGD_VG 2   // rotation parameters of the V kernel read this many slots at a time
#endif
#ifndef WN_EIGD_VWAVES
#define WN_EIGD_VWAVES 2   // waves per SIMD the V kernel's registers are sized for
#endif
#ifndef WN_EIGD_VDEPTH
#define WN_EIGD_VDEPTH 8   // record steps the V kernel loads ahead (divides the record rounds per sweep)
#endif

template <int N>
struct WnEigd {
    static constexpr int HALF = N / 2;
    static constexpr int NB = WnSlots<N>::NB;
    static constexpr int NOFF = HALF * (HALF - 1) / 2;
    static constexpr int LPK = N;                      // lanes per point in the sweep kernel
    static constexpr int KPW = 64 / LPK;               // points per wave in the sweep kernel
    static constexpr int KPV = 64 / N;                 // points per wave in the V kernel (one lane per row)
    static constexpr int RPS = N;                      // record rounds per sweep (N - 1 used, padded to N)
    // per point: records [sweep][RPS][N/2] double4 (2 double2 each), saved H [4 NB], state (4 double2).
    // RPS is a multiple of 4, so the V kernel's four-deep record pipeline has the same ring slots every sweep.
    static constexpr size_t REC2 = (size_t)WN_EIGD_SWEEPS * RPS * HALF * 2;
    static constexpr size_t PER_K = REC2 + 4 * NB + 4;
};

struct WnEigdState {
    int sweeps;        // sweeps done
    int nrec;          // sweeps recorded for the V kernel
    int cont;          // 1: finished by k_eig_rv_cont from the saved H
    int pad;
    double prev_off;   // off-diagonal mass at the stopping test before the last one (continuation)
    unsigned char rank[32];
};
static_assert(sizeof(WnEigdState) <= 64, "state slot");

struct WnEigdScratch {
    double2 *rec;           // [count][REC2]
    double2 *hsave;         // [count][4 NB]
    WnEigdState *state;     // [count]
};

template <int N>
static WnEigdScratch wn_eigd_layout(double2 *scr, int count) {
    using D = WnEigd<N>;
    WnEigdScratch s;
    s.rec = scr;
    s.hsave = scr + (size_t)count * D::REC2;
    s.state = (WnEigdState *)(s.hsave + (size_t)count * 4 * D::NB);
    return s;
}

// Point of item b of a pass: pass < 0 and pass 2 take b itself; passes 0 and 1 the warm-start groups.
__device__ __forceinline__ int wn_eig_item_k(int b, int items, int pass, int count, int *anchor) {
    *anchor = -1;
    if (b >= items) return -1;
    if (pass < 0 || pass == 2) return b;
    return wn_warm_k(b, pass, count, anchor);
}

// LDS layout of one point's H in the sweep kernel (lane-major).  Each lane owns up to four off-diagonal 2x2
// blocks: a 2x2 group of the slot grid (rows 2a, 2a+1 x columns 2b, 2b+1, a < b; lanes l < NGF, row-major over
// the groups) or two of the blocks (2q, 2q + 1) next to the diagonal (lanes NGF + q / 2).  Element e of the
// lane's block i sits at (4 i + e) LPK + lane, so the lanes of one access touch consecutive 16-byte words; the
// diagonal blocks' entries (a, h, d) sit in three planes of HALF words after them.  A round moves every element
// (x, y) to (sigma(x), sigma(y)), mostly the neighbouring group's slot, so the writes are runs of consecutive
// words.  At N = 32 the kernel places the elements by WnLaneWord instead (same buffer size).
template <int N>
struct WnLane {
    static constexpr int HALF = N / 2, G = HALF / 2, NGF = G * (G - 1) / 2, LPK = N;
    static constexpr int DIAG = 16 * LPK;
    static constexpr int SIZE = DIAG + 3 * HALF;   // double2 words per point
    static_assert(NGF + (G + 1) / 2 <= LPK, "block groups per point");
    // storage word of element (x, y), x <= y (slot positions)
    __host__ __device__ static constexpr int at(int x, int y) {
        const int t1 = x >> 1, t2 = y >> 1;
        if (t1 == t2) return DIAG + (x == y ? ((x & 1) ? 2 : 0) : 1) * HALF + t1;
        const int a = t1 >> 1, b = t2 >> 1;
        int lane = 0, i = 0;
        if (a < b) {
            lane = a * (G - 1) - a * (a - 1) / 2 + (b - a - 1);
            i = (t1 & 1) * 2 + (t2 & 1);
        } else {   // a == b: the block (2a, 2a + 1) next to the diagonal, the second of lane NGF + a / 2 if a is odd
            lane = NGF + (a >> 1);
            i = (a & 1) ? 3 : 0;
        }
        return (i * 4 + (x & 1) * 2 + (y & 1)) * LPK + lane;
    }
};

// Word of element (x, y), x <= y, in k_eigd_sweeps' buffer: WnLane's layout, or at N = 32 a placement that
// makes every 16-byte LDS read and write of a round free of bank conflicts on gfx950 (reads are served 16
// lanes per cycle from 64 banks, writes 8 lanes per cycle from 32 banks; WnLane's writes land on equal banks
// within a cycle's lanes for about half their words).  The placement is a table (gen_eigd_lds32.py).
// A lane's blocks 1 and 2 that it does not hold (lanes next to the diagonal) read and write spare words after
// the elements (SPARE of them), so that every lane runs the same instructions: spare(l, i, e) gives the read
// word (low 16 bits) and the write word (high 16 bits).
template <int N>
struct WnLaneWord {
    static constexpr int SPARE = 16;
    static constexpr int WORDS = WnLane<N>::SIZE + SPARE;
    __device__ static int at(int x, int y) { return WnLane<N>::at(x, y); }
    __device__ static unsigned spare(int, int i, int e) {
        const unsigned w = WnLane<N>::SIZE + ((4 * i + e) & (SPARE - 1));
        return w | (w << 16);
    }
};
template <>
struct WnLaneWord<32> {
    static constexpr int SPARE = 16;
    static constexpr int WORDS = WnLane<32>::SIZE + SPARE;
    __device__ static int at(int x, int y) { return (int)wn_eigd32_word[x * 32 - x * (x - 1) / 2 + (y - x)]; }
    // spare words of free bank classes in the lane's read and write cycles (lanes 28..31, blocks 1 and 2)
    __device__ static unsigned spare(int l, int i, int e) { return wn_eigd32_spare[(l - 28) & 3][(i - 1) & 1][e]; }
};
// destination of the element now at (x, y), x < y, after the round: word, bit 31 set when stored conjugated
template <int N>
__device__ __forceinline__ int wn_lane_dest(int x, int y) {
    const int sx = WnSlots<N>::sigma(x), sy = WnSlots<N>::sigma(y);
    return sx < sy ? WnLaneWord<N>::at(sx, sy) : (WnLaneWord<N>::at(sy, sx) | (int)0x80000000);
}

// 2x2 block groups of lane l of a point in k_eigd_sweeps: block i uses row slot rsl[i >> 1] and column slot
// csl[i & 1]; lanes l < NGF hold a group of four blocks, the others up to two blocks next to the diagonal
template <int N>
__device__ __forceinline__ void wn_lane_blocks(int l, int (&rsl)[2], int (&csl)[2], bool (&bon)[4]) {
    constexpr int G = WnLane<N>::G, NGF = WnLane<N>::NGF;
    rsl[0] = rsl[1] = csl[0] = csl[1] = 0;
    bon[0] = bon[1] = bon[2] = bon[3] = false;
    if (l < NGF) {
        int a = 0, rem = l;
        while (rem >= G - 1 - a) {
            rem -= G - 1 - a;
            ++a;
        }
        const int b = a + 1 + rem;
        rsl[0] = 2 * a;
        rsl[1] = 2 * a + 1;
        csl[0] = 2 * b;
        csl[1] = 2 * b + 1;
        bon[0] = bon[1] = bon[2] = bon[3] = true;
    } else {
        const int q0 = 2 * (l - NGF), q1 = q0 + 1;
        if (q0 < G) {
            rsl[0] = 2 * q0;
            csl[0] = 2 * q0 + 1;
            bon[0] = true;
        }
        if (q1 < G) {
            rsl[1] = 2 * q1;
            csl[1] = 2 * q1 + 1;
            bon[3] = true;
        }
    }
}

template <int N>
// two waves per SIMD (at N = 32 the LDS holds no more: two points per wave, 19 KB per wave)
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(2))) void
k_eigd_sweeps(const double *__restrict__ x, double *__restrict__ eout, int *__restrict__ sweeps_out,
              unsigned long long *__restrict__ sweep_total, int n, int ldo, double tol, int pass, int count,
              int items, WnEigdScratch scr, int ctr_stride, int *__restrict__ status, int max_sweeps,
              int rec_sweeps) {
    using D = WnEigd<N>;
    using S = WnSlots<N>;
    using L = WnLane<N>;
    using W = WnLaneWord<N>;
    constexpr int HALF = D::HALF, NB = D::NB, LPK = D::LPK, KPW = D::KPW;
    __shared__ double2 Hs[KPW][W::WORDS];
    // the round's parameters as two planes, (c, s) at slots 0 .. HALF - 1 and (cos phi, sin phi) after them:
    // lanes reading consecutive slots then touch consecutive 16-byte words
    __shared__ double2 Rp[KPW][N];
    const int lane = threadIdx.x;
    const int j = lane / LPK, l = lane - (lane / LPK) * LPK;
    const int item = blockIdx.x * KPW + j;
    int anchor;
    const int k = wn_eig_item_k(item, items, pass, count, &anchor);
    const bool active = k >= 0 && k < count;
    // H(k) upper triangle in band order -> slot positions of round 0 (unused words zero)
    for (int e = l; e < W::WORDS; e += LPK) Hs[j][e] = make_double2(0.0, 0.0);
    wn_wave_order();
    const double2 *xk = active ? (const double2 *)(x + (size_t)k * ldo) : nullptr;
    for (int e = l; e < N * N; e += LPK) {
        const int i = e / N, jj = e - (e / N) * N;
        if (i <= jj) {
            double2 hv = make_double2(0.0, 0.0);
            if (active && jj < n) {
                hv = xk[i * n + jj];
                if (i == jj) hv.y = 0.0;
            }
            const int pi = S::pos0(i), pj = S::pos0(jj);
            if (pi <= pj)
                Hs[j][W::at(pi, pj)] = hv;
            else
                Hs[j][W::at(pj, pi)] = make_double2(hv.x, -hv.y);
        }
    }
    // this lane's blocks, the words it reads them from and the words their rotated elements go to (spare words
    // for blocks it does not hold: their results are discarded)
    int rsl[2], csl[2];
    bool bon[4];
    wn_lane_blocks<N>(l, rsl, csl, bon);
    int brd[4][4], bdst[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int t1 = rsl[i >> 1], t2 = csl[i & 1];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int x = 2 * t1 + (e >> 1), y = 2 * t2 + (e & 1);
            const unsigned sp = bon[i] ? 0u : W::spare(l, i, e);
            brd[i][e] = bon[i] ? W::at(x, y) : (int)(sp & 0xffffu);
            bdst[i][e] = bon[i] ? wn_lane_dest<N>(x, y) : (int)(sp >> 16);
        }
    }
    // slot cs = l mod HALF: lanes l and l + HALF read its diagonal block and compute its rotation alike; lanes
    // l < HALF store (c, s) and the new a, the others (cos phi, sin phi) and the new d
    const bool plane = l < HALF;
    const int cs = l & (HALF - 1);
    const int rd_a = W::at(2 * cs, 2 * cs), rd_h = W::at(2 * cs, 2 * cs + 1), rd_d = W::at(2 * cs + 1, 2 * cs + 1);
    const int dst_ad = plane ? W::at(S::sigma(2 * cs), S::sigma(2 * cs)) : W::at(S::sigma(2 * cs + 1), S::sigma(2 * cs + 1));
    const int dst_h = wn_lane_dest<N>(2 * cs, 2 * cs + 1) & 0x7fffffff;   // never conjugated (sigma keeps the order)
    const int rec_at = 2 * cs + (plane ? 0 : 1);                           // this lane's half of the slot's record
    // words of the stopping test's blocks b = l, l + N, ... (k_eig_rv's packed slot order): (2 t1, 2 t2),
    // (2 t1, 2 t2 + 1), (2 t1 + 1, 2 t2 + 1), (2 t1 + 1, 2 t2); the last is unused on the diagonal (rem == 0)
    constexpr int NST = (NB + N - 1) / N;
    int stw[NST][4];
    bool stdiag[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int b = min(l + q * N, NB - 1);
        int t1 = 0, rem = b;
        while (rem >= HALF - t1) {
            rem -= HALF - t1;
            ++t1;
        }
        const int t2 = t1 + rem;
        stdiag[q] = rem == 0;
        stw[q][0] = W::at(2 * t1, 2 * t2);
        stw[q][1] = W::at(2 * t1, 2 * t2 + 1);
        stw[q][2] = W::at(2 * t1 + 1, 2 * t2 + 1);
        stw[q][3] = rem == 0 ? stw[q][0] : W::at(2 * t1 + 1, 2 * t2);
    }
    // records of this point; a point past the last item writes (identity rotations, padding round only) into
    // the records of the wave's first point, which is active whenever any point of the wave is
    double2 *rec = scr.rec + (size_t)(active ? item : (int)blockIdx.x * KPW) * D::REC2;
    wn_wave_order();

    int sweep = 0;
    bool frozen = !active, converged = false, cont = false;
    double prev_off = 0.0, saved_prev = 0.0;
    for (;;) {
        // stopping test of k_eig_rv, with its summation order (blocks b = l, l + N, ... of the packed slot
        // order, then the lane tree): ||offdiag||_F against tol ||H||_F, or the rounding floor
        double off = 0.0, dia = 0.0;
        if (!frozen) {
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                if (l + q * N >= NB) break;
                const double2 e0 = Hs[j][stw[q][0]], e1 = Hs[j][stw[q][1]];
                const double2 e3 = Hs[j][stw[q][2]];
                const double m0 = e0.x * e0.x + e0.y * e0.y, m1 = e1.x * e1.x + e1.y * e1.y;
                const double m3 = e3.x * e3.x + e3.y * e3.y;
                if (stdiag[q]) {
                    dia += m0 + m3;
                    off += 2.0 * m1;
                } else {
                    const double2 e2 = Hs[j][stw[q][3]];
                    off += 2.0 * (m0 + m1 + m3 + e2.x * e2.x + e2.y * e2.y);
                }
            }
        }
#pragma unroll
        for (int o = N / 2; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            dia += __shfl_xor(dia, o);
        }
        if (!frozen) {
            if (off <= tol * tol * (dia + off) ||
                (sweep > 0 && off > 0.25 * prev_off && off <= 1e-20 * (dia + off))) {
                frozen = true;
                converged = true;
            } else if (sweep >= max_sweeps) {
                frozen = true;   // sweep cap: reported through status word 1
            } else if (sweep >= rec_sweeps) {
                frozen = true;   // finished by k_eig_rv_cont from this state
                cont = true;
                saved_prev = prev_off;
            }
        }
        prev_off = off;
        if (__ballot(!frozen) == 0ull) break;
        // a frozen point runs the rounds with identity rotations (its H only moves through the slots and is
        // back in place at the end of the sweep) and writes its records into the padding round, which the V
        // kernel never applies
        double2 *rs = rec + (size_t)min(sweep, WN_EIGD_SWEEPS - 1) * D::RPS * N;
        for (int r = 0; r < N - 1; ++r) {
            // every read of the round first
            double2 B[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e) B[i][e] = Hs[j][brd[i][e]];
            const double a = Hs[j][rd_a].x;
            const double2 h = Hs[j][rd_h];
            const double d = Hs[j][rd_d].x;
            const int2 pq = wn_rr_rt(N, cs, r);
            const WnRotParams o = wn_rot_params(a, d, h, !frozen && pq.x < n && pq.y < n);
            const double2 half = plane ? make_double2(o.R.x, o.R.y) : make_double2(o.R.z, o.R.w);
            Rp[j][l] = half;
            rs[(size_t)(frozen ? N - 1 : r) * N + rec_at] = half;
            wn_wave_order();
            {
                double4 Rr[2], Rc[2];   // rotations of the row slots and the column slots
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const double2 a0 = Rp[j][rsl[hh]], a1 = Rp[j][HALF + rsl[hh]];
                    const double2 b0 = Rp[j][csl[hh]], b1 = Rp[j][HALF + csl[hh]];
                    Rr[hh] = make_double4(a0.x, a0.y, a1.x, a1.y);
                    Rc[hh] = make_double4(b0.x, b0.y, b1.x, b1.y);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double4 R1 = Rr[i >> 1], R2 = Rc[i & 1];
                    double2 M[2][2] = {{B[i][0], B[i][1]}, {B[i][2], B[i][3]}};
                    // B J2, then J1^+ (B J2): the block update of wn_block_update
                    const double2 em2 = make_double2(R2.z, -R2.w);
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const double2 y = cmul(em2, M[u][1]);
                        const double2 x0 = M[u][0];
                        M[u][0] = make_double2(R2.x * x0.x - R2.y * y.x, R2.x * x0.y - R2.y * y.y);
                        M[u][1] = make_double2(R2.y * x0.x + R2.x * y.x, R2.y * x0.y + R2.x * y.y);
                    }
                    const double2 ep1 = make_double2(R1.z, R1.w);
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        const double2 y = cmul(ep1, M[1][v]);
                        const double2 x0 = M[0][v];
                        M[0][v] = make_double2(R1.x * x0.x - R1.y * y.x, R1.x * x0.y - R1.y * y.y);
                        M[1][v] = make_double2(R1.y * x0.x + R1.x * y.x, R1.y * x0.y + R1.x * y.y);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int dd = bdst[i][e];
                        const double2 val = M[e >> 1][e & 1];
                        Hs[j][dd & 0x7fffffff] =
                            make_double2(val.x, __longlong_as_double(__double_as_longlong(val.y) ^
                                                                     ((unsigned long long)(dd < 0) << 63)));
                    }
                }
                // the slot's new diagonal entries (a from lanes l < HALF, d from the others) and its new (left,
                // right) entry (lanes l and l + HALF write the same value)
                Hs[j][dst_ad] = make_double2(plane ? o.an : o.dn, 0.0);
                Hs[j][dst_h] = o.h;
            }
            wn_wave_order();
        }
        if (!frozen) ++sweep;
    }
    if (!active) return;
    WnEigdState *stp = scr.state + item;
    if (cont) {
        // H at a sweep boundary (every band at pos0) in k_eig_rv's packed slot order, for k_eig_rv_cont
        double2 *hs = scr.hsave + (size_t)item * 4 * NB;
        for (int e = l; e < 4 * NB; e += LPK) {
            const int pl = e / NB, b = e - (e / NB) * NB;
            int t1 = 0, rem = b;
            while (rem >= HALF - t1) {
                rem -= HALF - t1;
                ++t1;
            }
            const int xx = 2 * t1 + (pl >> 1), yy = 2 * (t1 + rem) + (pl & 1);
            hs[e] = xx <= yy ? Hs[j][W::at(xx, yy)] : make_double2(0.0, 0.0);   // (1, 0) of a diagonal block: unused
        }
        if (l == 0) {
            stp->sweeps = sweep;
            stp->nrec = sweep;
            stp->cont = 1;
            stp->prev_off = saved_prev;
        }
        return;
    }
    // eigenvalues ascending (stable by index) and their ranks
    if (l < n) {
        auto diag = [&](int b) {
            const int p = S::pos0(b);
            return Hs[j][W::at(p, p)].x;
        };
        const double et = diag(l);
        int rk = 0;
        for (int jj = 0; jj < n; ++jj) {
            const double ej = diag(jj);
            rk += (ej < et) || (ej == et && jj < l);
        }
        stp->rank[l] = (unsigned char)rk;
        eout[(size_t)k * n + rk] = et;
    }
    if (l == 0) {
        stp->sweeps = sweep;
        stp->nrec = sweep;
        stp->cont = 0;
        const int bucket = ctr_stride * (int)(item % WN_CTR_BUCKETS);
        if (sweeps_out) atomicMax(sweeps_out + bucket, sweep);
        if (sweep_total) atomicAdd(sweep_total + bucket / 2, (unsigned long long)sweep);
        if (!converged && status) atomicOr(status + 1, 1);
    }
}

// Step T of a sweep in the V kernel (T = 0 .. RPS - 1; rounds T < N - 1): pre[T % 4] holds this lane's share
// of round T's records (loaded four steps earlier); it is staged in LDS, the ring slot is refilled with the
// records of step T + 4, and the round is applied to the row.  Step N - 1 .. RPS - 1 only keep the pipeline
// moving.  g is the point's first record step of this sweep; later sweeps continue at g + RPS.
template <int N, int T>
struct WnVApply {
    __device__ __forceinline__ static void run(WnRow<N> &V, double4 (*stage)[N / 2], double2 (&pre)[WN_EIGD_VDEPTH],
                                               const double2 *rp, int g, int glim, int l) {
        // rp is not __restrict__: the wave-order fences must keep each record load at its step (a restrict
        // pointer lets the compiler hoist every round's load to the start and spill the row)
        // branch-free: the ring slot is refilled from a clamped step of a valid record region, and steps past
        // the point's own records are replaced by the identity when they are staged
        constexpr int GMAX = WN_EIGD_SWEEPS * WnEigd<N>::RPS;
        constexpr int DP = WN_EIGD_VDEPTH;
        static_assert(WnEigd<N>::RPS % DP == 0, "ring slots repeat every sweep");
        if constexpr (T < N - 1)
            ((double2 *)stage[T & 1])[l] = g + T < glim ? pre[T % DP] : make_double2(1.0, 0.0);
        pre[T % DP] = rp[(size_t)min(g + T + DP, GMAX - 1) * N + l];
        if constexpr (T < N - 1) {
            wn_wave_order();
            WnVSlots<N, T, 0, WN_EIGD_VG>::run(V, stage[T & 1]);
            wn_wave_order();
        }
        WnVApply<N, T + 1>::run(V, stage, pre, rp, g, glim, l);
    }
};
template <int N>
struct WnVApply<N, N> {   // N == WnEigd<N>::RPS
    __device__ __forceinline__ static void run(WnRow<N> &, double4 (*)[N / 2], double2 (&)[WN_EIGD_VDEPTH], const double2 *, int,
                                               int, int) {}
};

template <int N>
__global__ __launch_bounds__(64, WN_EIGD_VWAVES) void k_eigd_vapply(double2 *__restrict__ uout, const double2 *__restrict__ uprev,
                                                    int n, int pass, int count, int items, WnEigdScratch scr) {
    using D = WnEigd<N>;
    constexpr int KPV = D::KPV;
    __shared__ double4 stage[KPV][2][N / 2];
    __shared__ int rank[KPV][N];
    const int lane = threadIdx.x;
    const int j = lane / N, l = lane - (lane / N) * N;
    const int item = blockIdx.x * KPV + j;
    int anchor;
    const int k = wn_eig_item_k(item, items, pass, count, &anchor);
    const bool active = k >= 0 && k < count;
    int nrec = 0, cont = 0;
    if (active) {
        nrec = scr.state[item].nrec;
        cont = scr.state[item].cont;
        rank[j][l] = scr.state[item].rank[l];
    }
    const double2 *ua = !active ? nullptr
                                : (pass == 2 ? uprev + (size_t)k * n * n
                                             : (anchor >= 0 ? uout + (size_t)anchor * n * n : nullptr));
    WnRow<N> V;
    WnRowIo<N, 0>::load(V, ua, n, l);
    int nmax = nrec;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    // record steps of this point: sweep s, round r at step s * RPS + r; steps past its own sweeps (and the
    // padding rounds) read as identity rotations, which leave the row exactly as it is
    const double2 *rp = scr.rec + (active ? (size_t)item * D::REC2 : 0);
    const int glim = nrec * D::RPS;
    double2 pre[WN_EIGD_VDEPTH];
#pragma unroll
    for (int q = 0; q < WN_EIGD_VDEPTH; ++q) pre[q] = rp[(size_t)q * N + l];
    for (int s = 0; s < nmax; ++s) WnVApply<N, 0>::run(V, stage[j], pre, rp, s * D::RPS, glim, l);
    wn_wave_order();
    if (!active) return;
    double2 *uk = uout + (size_t)k * n * n;
    if (cont) {
        // raw rows for k_eig_rv_cont
        if (l < n) WnRowIo<N, 0>::store_raw(V, uk, n, l);
        return;
    }
    WnRowIo<N, 0>::store(V, uk, rank[j], n, l, j * N);
}

// Points that k_eigd_sweeps left unfinished: the register-V kernel's roles from the saved H, sweep count and
// stopping-test history, V from the rows the V kernel wrote.  One point per workgroup iteration.
template <int N>
__global__ __launch_bounds__(128, 2) void k_eig_rv_cont(WnEigdScratch scr, double *__restrict__ eout,
                                                        double2 *__restrict__ uout, int *__restrict__ sweeps_out,
                                                        unsigned long long *__restrict__ sweep_total, int n,
                                                        double tol, int pass, int count, int items, int ctr_stride,
                                                        int *__restrict__ status, int max_sweeps) {
    __shared__ WnEigShared<N> sm;
    __shared__ int list[128];
    __shared__ int nlist, ncnt0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = lane / N, l = lane - (lane / N) * N;
    // the continuation flags of 128 items at a time, read by one thread each and compacted in item order
    for (int base = blockIdx.x * 128; base < items; base += gridDim.x * 128) {
        const int it0 = base + (int)threadIdx.x;
        int anchor0;
        const int k0 = wn_eig_item_k(it0, items, pass, count, &anchor0);
        const bool want = k0 >= 0 && k0 < count && scr.state[it0].cont;
        const unsigned long long bal = __ballot(want);
        if (threadIdx.x == 0) ncnt0 = __popcll(bal);   // thread 0 is in wave 0
        __syncthreads();
        if (want) list[(wave == 0 ? 0 : ncnt0) + __popcll(bal & ((1ull << lane) - 1))] = it0;
        if (threadIdx.x == 64) nlist = ncnt0 + __popcll(bal);   // lane 0 of wave 1
        __syncthreads();
        const int nl = nlist;
        for (int c = 0; c < nl; ++c) {
            const int it = list[c];
            int anchor;
            const int k = wn_eig_item_k(it, items, pass, count, &anchor);
            const WnEigdState st = scr.state[it];
            const bool act = j == 0;
            if (wave == 0) {
                const int bucket = ctr_stride * (int)(it % WN_CTR_BUCKETS);
                wn_eig_h_role<N>(sm, nullptr, eout, sweeps_out, sweep_total, n, 0, tol, k, act, j, l, lane, bucket,
                                 status, max_sweeps, scr.hsave + (size_t)it * 4 * WnSlots<N>::NB, st.sweeps,
                                 st.prev_off);
            } else {
                wn_eig_v_role<N>(sm, uout, act ? uout + (size_t)k * n * n : nullptr, n, k, act, j, l, st.sweeps);
            }
        }
        __syncthreads();
    }
}

size_t wn_eigd_scratch_elems(int n, int count) {
    if (n < 1 || n > 32) return 0;
    const size_t per = n <= 8 ? WnEigd<8>::PER_K : n <= 16 ? WnEigd<16>::PER_K : WnEigd<32>::PER_K;
    return per * (size_t)count;
}

template <int N>
static hipError_t launch_eigd(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                              int count, const WnShape &sh, int pass, hipStream_t st, const double2 *uprev,
                              int ctr_stride, int *status, double2 *scr, int items) {
    using D = WnEigd<N>;
    const double tol = 1.0e-15;
    const WnEigdScratch s = wn_eigd_layout<N>(scr, items);
    hipLaunchKernelGGL((k_eigd_sweeps<N>), dim3((items + D::KPW - 1) / D::KPW), dim3(64), 0, st, x, e, sweeps,
                       sweep_total, sh.n, sh.ldo, tol, pass, count, items, s, ctr_stride, status, sh.max_sweeps,
                       std::min(std::max(sh.rec_sweeps, 1), WN_EIGD_SWEEPS));
    hipLaunchKernelGGL((k_eigd_vapply<N>), dim3((items + D::KPV - 1) / D::KPV), dim3(64), 0, st, u, uprev, sh.n,
                       pass, count, items, s);
    if (sh.max_sweeps > std::min(std::max(sh.rec_sweeps, 1), WN_EIGD_SWEEPS))
        hipLaunchKernelGGL((k_eig_rv_cont<N>), dim3(std::min((items + 127) / 128, 256)), dim3(128), 0, st, s, e, u,
                           sweeps,
                           sweep_total, sh.n, tol, pass, count, items, ctr_stride, status, sh.max_sweeps);
    return hipGetLastError();
}

hipError_t wn_launch_eig_defer(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                               int count, const WnShape &sh, int pass, hipStream_t st, const double2 *uprev,
                               int ctr_stride, int *status, double2 *scr) {
    const int n = sh.n;
    if (!wn_eig_rv_supported(n) || sh.max_sweeps < 1 || !scr) return hipErrorInvalidValue;
    if (pass == 2 && !uprev) return hipErrorInvalidValue;
    constexpr int G = 4;
    const int groups = (count + G - 1) / G;
    const int items = (pass < 0 || pass == 2) ? count : pass == 0 ? groups : groups * (G - 1);
    if (items <= 0) return hipSuccess;
    if (n <= 8) return launch_eigd<8>(x, e, u, sweeps, sweep_total, count, sh, pass, st, uprev, ctr_stride, status, scr,
                                      items);
    if (n <= 16)
        return launch_eigd<16>(x, e, u, sweeps, sweep_total, count, sh, pass, st, uprev, ctr_stride, status, scr,
                               items);
    return launch_eigd<32>(x, e, u, sweeps, sweep_total, count, sh, pass, st, uprev, ctr_stride, status, scr, items);
}

bool wn_eig_rv_supported(int n) { return n >= 1 && n <= 32; }

hipError_t wn_launch_eig_rv(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                            int count, const WnShape &sh, int pass, hipStream_t st, const double2 *uprev,
                            int ctr_stride, int *status) {
    const int n = sh.n;
    if (!wn_eig_rv_supported(n) || sh.max_sweeps < 1) return hipErrorInvalidValue;
    if (pass == 2 && !uprev) return hipErrorInvalidValue;
    const double tol = 1.0e-15;
    constexpr int G = 4;
    const int groups = (count + G - 1) / G;
    const int items = (pass < 0 || pass == 2) ? count : pass == 0 ? groups : groups * (G - 1);
    if (items <= 0) return hipSuccess;
#define WN_RV_LAUNCH(NN)                                                                                          \
    hipLaunchKernelGGL((k_eig_rv<NN>), dim3((items + 64 / NN - 1) / (64 / NN)), dim3(128), 0, st, x, e, u, sweeps, \
                       sweep_total, n, sh.ldo, tol, pass, count, items, uprev, ctr_stride, status, sh.max_sweeps)
    if (n <= 8)
        WN_RV_LAUNCH(8);
    else if (n <= 16)
        WN_RV_LAUNCH(16);
    else
        WN_RV_LAUNCH(32);
#undef WN_RV_LAUNCH
    return hipGetLastError();
}
