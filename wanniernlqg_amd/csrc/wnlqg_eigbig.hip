// Batched Hermitian Jacobi eigensolver for 64 < n_w <= 128 with deferred eigenvectors.
//
// Same mathematics and sweep order as the LDS kernel k_eig_f64 (wnlqg_kernels.hip): the upper triangle
// of channel 0 is the Hermitian matrix (DenseSpectra.hermitian_spectrum, reference
// src/Core/ESKernel/DenseSpectra.jl:14-16, uplo 'U'), parallel cyclic Jacobi over a round-robin
// tournament (element 0 fixed, the others moving one position per round), the 2x2 block form of a round
// on the packed upper triangle, the same stopping rule, ascending sort and canonical column phase
// (canonicalize_eigenvectors!, HamiltonianMatrixElements.jl:6-25).
//
// At these sizes H (packed, 74.5 KB at N = 96) and V (147 KB) do not fit the LDS together, and updating
// V in global memory inside the sweep loop puts a global-memory round trip and a fence into every round.
// Here the work is split by what it depends on:
//   * k_eigb_sweeps: one workgroup per k point runs up to WN_EIGB_SB sweeps on H alone (H in LDS) and
//     records every round's rotations (c, s, cos phi, sin phi) in global memory; H and the solver state
//     go back to global memory between launches.
//   * k_eigb_vapply: V <- V J_1 J_2 ... for the recorded rounds.  Row i of V only ever combines entries
//     of row i, so a wave owns RW whole rows (in LDS) and applies every round to them without any
//     barrier: the LDS executes one wave's accesses in program order.  The rotation records of a k point
//     are read by all of its row waves (run on one XCD, so that they come from that XCD's L2).
//   * k_eigb_finish: ascending sort, canonical phase, eigenvectors and counters.
// The host alternates the first two for ceil(max_sweeps / WN_EIGB_SB) chunks; a k point that has
// converged skips the remaining chunks at once.
//
// Scratch per k point (complex elements): V [N][N] row-major, the packed H, the rotation records of one
// chunk, the diagonal at convergence and the state (sweeps done, done, rounds recorded, converged; the
// previous off-diagonal norm).

#include "wnlqg_common.h"
#include "wnlqg_kernels.h"

#define WN_EIGB_SB 4   // sweeps per chunk launch
#define WN_EIGB_RW 4   // V rows per wave

namespace {

template <int N>
struct EigbLayout {
    static constexpr int HALF = N / 2;
    static constexpr int NP = N * (N + 1) / 2;
    static constexpr size_t HG = (size_t)N * N;                                        // packed H
    static constexpr size_t REC = HG + NP;                                             // double4 records
    static constexpr size_t DIAG = REC + (size_t)WN_EIGB_SB * (N - 1) * HALF * 2;      // N doubles
    static constexpr size_t ST = DIAG + (N + 1) / 2;                                  // int4 + double
    static constexpr size_t PER_K = ST + 2;
};

// packed row-major index of (i, j), i <= j
template <int N>
__device__ __forceinline__ int eigb_p(int i, int j) {
    return (int)(__umul24((unsigned)i, (unsigned)(2 * N - i + 1)) >> 1) + (j - i);
}

// k point of slot s of a pass: pass -1 / 2 every point; pass 0 the warm-start group anchors; pass 1 the other
// points of each group (anchor set); a slot past the last point of a short group maps to k >= count.
__device__ __forceinline__ int eigb_point(int s, int pass, int count, int *anchor) {
    if (pass == 0 || pass == 1) return wn_warm_k(s, pass, count, anchor);
    *anchor = -1;
    return s;
}

int eigb_slots(int pass, int count) {
    const int groups = (count + WN_WARM_G - 1) / WN_WARM_G;
    return pass == 0 ? groups : (pass == 1 ? groups * (WN_WARM_G - 1) : count);
}

}  // namespace

template <int N, int TPB>
__global__ __launch_bounds__(TPB, 1) void k_eigb_sweeps(const double *__restrict__ x, int n, int ldo, double tol,
                                                        int count, double2 *__restrict__ scr, int chunk,
                                                        int max_sweeps, int pass) {
    using L = EigbLayout<N>;
    constexpr int HALF = L::HALF, NP = L::NP, NW = TPB / 64;
    __shared__ double2 H[NP];
    __shared__ double4 rot[HALF];   // (c, s, cos phi, sin phi)
    __shared__ int2 pq[HALF];
    __shared__ double red[2 * NW];
    int anchor;
    const int k = eigb_point((int)blockIdx.x, pass, count, &anchor);
    if (k >= count) return;
    const int t = threadIdx.x;
    double2 *sk = scr + (size_t)k * L::PER_K;
    int4 *state = (int4 *)(sk + L::ST);
    double *prevp = (double *)(sk + L::ST + 1);
    int4 stv = chunk == 0 ? make_int4(0, 0, 0, 0) : *state;   // (sweeps, done, rounds recorded, converged)
    if (stv.y) {
        if (t == 0) state->z = 0;
        return;
    }
    double prev_off = chunk == 0 ? 0.0 : *prevp;
    double2 *hg = sk + L::HG;
    double4 *rec = (double4 *)(sk + L::REC);
    if (chunk == 0) {
        const double2 *xk = (const double2 *)(x + (size_t)k * ldo);
        for (int e = t; e < N * N; e += TPB) {
            const int i = e / N, j = e - (e / N) * N;
            if (i <= j) {
                double2 hv = make_double2(0.0, 0.0);
                if (j < n) {
                    hv = xk[i * n + j];
                    if (i == j) hv.y = 0.0;
                }
                H[eigb_p<N>(i, j)] = hv;
            }
        }
    } else {
        for (int e = t; e < NP; e += TPB) H[e] = hg[e];
    }
    __syncthreads();

    // off-diagonal blocks (t1 < t2) of a round, a fixed share per thread
    constexpr int NOFF = HALF * (HALF - 1) / 2;
    constexpr int BPT = (NOFF + TPB - 1) / TPB;
    int bt1[BPT], bt2[BPT];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int bidx = t + q * TPB;
        int t1 = 0, rem = bidx;
        while (t1 < HALF - 1 && rem >= HALF - 1 - t1) {
            rem -= HALF - 1 - t1;
            ++t1;
        }
        const bool ok = bidx < NOFF;
        bt1[q] = ok ? t1 : -1;
        bt2[q] = ok ? t1 + 1 + rem : -1;
    }

    int done_sweeps = 0;
    bool converged = false;
    while (done_sweeps < WN_EIGB_SB && stv.x < max_sweeps) {
        double off = 0.0, dia = 0.0;
        for (int e = t; e < N * N; e += TPB) {
            const int i = e / N, j = e - (e / N) * N;
            if (i > j) continue;
            const double2 hv = H[eigb_p<N>(i, j)];
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
        // converged, or at the rounding floor (a sweep no longer reduces the off-diagonal mass
        // quadratically once it is below ~1e-20 of the total)
        if (off <= tol * tol * (dia + off) || (stv.x > 0 && off > 0.25 * prev_off && off <= 1e-20 * (dia + off))) {
            converged = true;
            break;
        }
        prev_off = off;
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
                    const double a = H[eigb_p<N>(p, p)].x, d = H[eigb_p<N>(q, q)].x;
                    const double2 h = H[eigb_p<N>(p, q)];
                    const double b2 = h.x * h.x + h.y * h.y;
                    if (b2 > 0.0) {
                        // t = tan(theta) = sign(d-a) 2b / (|d-a| + sqrt((d-a)^2 + 4b^2))
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
                        H[eigb_p<N>(p, p)] = make_double2(a - tt * b, 0.0);
                        H[eigb_p<N>(q, q)] = make_double2(d + tt * b, 0.0);
                        H[eigb_p<N>(p, q)] = make_double2(0.0, 0.0);
                    }
                }
                const double4 R = make_double4(c, s, cph, sph);
                rot[t] = R;
                pq[t] = make_int2(p, q);
                rec[((size_t)done_sweeps * (N - 1) + r) * HALF + t] = R;
            }
            __syncthreads();
#pragma unroll
            for (int qb = 0; qb < BPT; ++qb) {
                const int t1 = bt1[qb], t2 = bt2[qb];
                if (t1 < 0) continue;
                const double4 R1 = rot[t1], R2 = rot[t2];
                if (R1.y == 0.0 && R2.y == 0.0) continue;
                const int2 a1 = pq[t1], a2 = pq[t2];
                const int rr[2] = {a1.x, a1.y}, cc[2] = {a2.x, a2.y};
                // element (i, j) is stored at (min, max) and conjugated when i > j
                double2 B[2][2];
                int offs[2][2];
                unsigned long long flip[2][2];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        const int i = rr[u], j = cc[v];
                        offs[u][v] = eigb_p<N>(min(i, j), max(i, j));
                        flip[u][v] = (unsigned long long)(i > j) << 63;
                    }
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        const double2 h = H[offs[u][v]];
                        B[u][v] = make_double2(h.x, __longlong_as_double(__double_as_longlong(h.y) ^ flip[u][v]));
                    }
                // B J2, then J1^+ (B J2)
                const double2 em2 = make_double2(R2.z, -R2.w);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const double2 y = cmul(em2, B[u][1]);
                    const double2 x0 = B[u][0];
                    B[u][0] = make_double2(R2.x * x0.x - R2.y * y.x, R2.x * x0.y - R2.y * y.y);
                    B[u][1] = make_double2(R2.y * x0.x + R2.x * y.x, R2.y * x0.y + R2.x * y.y);
                }
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
                        H[offs[u][v]] = make_double2(B[u][v].x, __longlong_as_double(__double_as_longlong(B[u][v].y) ^
                                                                                     flip[u][v]));
            }
            __syncthreads();
        }
        ++done_sweeps;
        ++stv.x;
    }
    const bool finished = converged || stv.x >= max_sweeps;
    if (finished) {
        double *diag = (double *)(sk + L::DIAG);
        for (int i = t; i < N; i += TPB) diag[i] = H[eigb_p<N>(i, i)].x;
    } else {
        for (int e = t; e < NP; e += TPB) hg[e] = H[e];
    }
    if (t == 0) {
        *state = make_int4(stv.x, finished ? 1 : 0, done_sweeps * (N - 1), converged ? 1 : 0);
        *prevp = prev_off;
    }
}

// V <- V J for the rounds recorded by the last sweep chunk.  Workgroup b of 4 waves: slot and row group from b
// (all row groups of one point on XCD b % 8); wave w owns rows (4 g + w) RW .. + RW - 1.  The first chunk starts
// V at the identity, or at the warm start's eigenvectors: the group anchor's (pass 1, from u) or the same
// point's of the previous mesh plane (pass 2, uprev).
template <int N, int RW>
__global__ __launch_bounds__(256) void k_eigb_vapply(int n, int count, double2 *__restrict__ scr, int chunk,
                                                     int groups, int pass, int nslots, const double2 *__restrict__ u,
                                                     const double2 *__restrict__ uprev) {
    using L = EigbLayout<N>;
    constexpr int HALF = L::HALF;
    constexpr int ITEMS = RW * HALF;
    constexpr int IPL = (ITEMS + 63) / 64;
    __shared__ double2 rows[4][RW][N];
    const int b = blockIdx.x, xcd = b & 7, s8 = b >> 3;
    const int slot = (s8 / groups) * 8 + xcd;
    if (slot >= nslots) return;
    int anchor;
    const int k = eigb_point(slot, pass, count, &anchor);
    if (k >= count) return;
    const int g = s8 - (s8 / groups) * groups;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = (4 * g + w) * RW;
    if (row0 >= n) return;   // whole wave; this kernel has no barrier
    double2 *sk = scr + (size_t)k * L::PER_K;
    const int nrec = ((const int4 *)(sk + L::ST))->z;
    if (nrec == 0) return;
    const double4 *rec = (const double4 *)(sk + L::REC);
    double2 *vg = sk;   // V row-major [N][N]
    double2 (*rw)[N] = rows[w];
    const double2 *w0 = pass == 1 ? u + (size_t)anchor * n * n : (pass == 2 ? uprev + (size_t)k * n * n : nullptr);
    for (int e = lane; e < RW * N; e += 64) {
        const int j = e / N, c = e - (e / N) * N, i = row0 + j;
        double2 v;
        if (chunk != 0)
            v = vg[(size_t)i * N + c];
        else if (w0)
            v = (i < n && c < n) ? w0[(size_t)i * n + c] : make_double2(0.0, 0.0);
        else
            v = make_double2(i == c ? 1.0 : 0.0, 0.0);
        rw[j][c] = v;
    }
    // items of this lane: (row j, pair slot s); the slot's pair follows the round-robin of the sweep kernel
    int js[IPL], ss[IPL], xp[IPL], xq[IPL];
    bool on[IPL];
#pragma unroll
    for (int u = 0; u < IPL; ++u) {
        const int e = lane + 64 * u;
        on[u] = e < ITEMS && row0 + e / HALF < n;
        js[u] = on[u] ? e / HALF : 0;
        ss[u] = on[u] ? e - (e / HALF) * HALF : 0;
    }
    double4 Rn[IPL];
#pragma unroll
    for (int u = 0; u < IPL; ++u) Rn[u] = rec[ss[u]];
    for (int r = 0; r < nrec; ++r) {
        const int rr = r % (N - 1);
        if (rr == 0) {
#pragma unroll
            for (int u = 0; u < IPL; ++u) {
                xp[u] = ss[u] >= 1 ? ss[u] - 1 : 0;
                xq[u] = N - 2 - ss[u];
            }
        }
        double4 R[IPL];
#pragma unroll
        for (int u = 0; u < IPL; ++u) R[u] = Rn[u];
        if (r + 1 < nrec) {
#pragma unroll
            for (int u = 0; u < IPL; ++u) Rn[u] = rec[(size_t)(r + 1) * HALF + ss[u]];
        }
#pragma unroll
        for (int u = 0; u < IPL; ++u) {
            int p = ss[u] == 0 ? 0 : xp[u] + 1;
            int q = xq[u] + 1;
            xp[u] = xp[u] == N - 2 ? 0 : xp[u] + 1;
            xq[u] = xq[u] == N - 2 ? 0 : xq[u] + 1;
            if (p > q) {
                const int tmp = p;
                p = q;
                q = tmp;
            }
            if (on[u]) {
                const double4 Q = R[u];
                const double2 xv = rw[js[u]][p], yv = rw[js[u]][q];
                const double2 ey = make_double2(Q.z * yv.x + Q.w * yv.y, Q.z * yv.y - Q.w * yv.x);
                rw[js[u]][p] = make_double2(Q.x * xv.x - Q.y * ey.x, Q.x * xv.y - Q.y * ey.y);
                rw[js[u]][q] = make_double2(Q.y * xv.x + Q.x * ey.x, Q.y * xv.y + Q.x * ey.y);
            }
        }
    }
    for (int e = lane; e < RW * N; e += 64) {
        const int j = e / N, c = e - (e / N) * N, i = row0 + j;
        if (i < n) vg[(size_t)i * N + c] = rw[j][c];
    }
}

// Ascending sort (stable by index), canonical phase per eigenvector column (first largest |v_i| made real
// and non-negative), eigenvalues and eigenvectors in sorted order, counters and the sweep-cap status.
template <int N>
__global__ __launch_bounds__(256) void k_eigb_finish(double *__restrict__ eout, double2 *__restrict__ uout,
                                                     int *__restrict__ sweeps_out,
                                                     unsigned long long *__restrict__ sweep_total, int n, int count,
                                                     const double2 *__restrict__ scr, int ctr_stride,
                                                     int *__restrict__ status, int pass) {
    using L = EigbLayout<N>;
    __shared__ double ev[N];
    __shared__ int rank[N];
    __shared__ double2 phase[N];
    int anchor;
    const int k = eigb_point((int)blockIdx.x, pass, count, &anchor);
    if (k >= count) return;
    const int t = threadIdx.x;
    const double2 *sk = scr + (size_t)k * L::PER_K;
    const double *diag = (const double *)(sk + L::DIAG);
    const int4 stv = *(const int4 *)(sk + L::ST);
    const double2 *vg = sk;
    if (t < n) ev[t] = diag[t];
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
        double pm = hypot(vg[t].x, vg[t].y);
        for (int i = 1; i < n; ++i) {
            const double2 v = vg[(size_t)i * N + t];
            const double m = hypot(v.x, v.y);
            if (m > pm) {
                pm = m;
                piv = i;
            }
        }
        double2 ph = make_double2(1.0, 0.0);
        if (pm > 0.0) {
            const double2 v = vg[(size_t)piv * N + t];
            ph = make_double2(v.x / pm, -v.y / pm);
        }
        phase[t] = ph;
    }
    __syncthreads();
    double2 *uk = uout + (size_t)k * n * n;
    for (int e = t; e < n * n; e += 256) {
        const int i = e / n, j = e - i * n;
        uk[i * n + rank[j]] = cmul(vg[(size_t)i * N + j], phase[j]);
    }
    const int bucket = ctr_stride * (int)(blockIdx.x % WN_CTR_BUCKETS);
    if (t == 0 && sweeps_out) atomicMax(sweeps_out + bucket, stv.x);
    if (t == 0 && sweep_total) atomicAdd(sweep_total + bucket / 2, (unsigned long long)stv.x);
    if (t == 0 && !stv.w && status) atomicOr(status + 1, 1);
}

size_t wn_eig_big_scratch_elems(int n, int count) {
    if (n <= 64) return 0;
    return (n <= 96 ? EigbLayout<96>::PER_K : EigbLayout<128>::PER_K) * (size_t)count;
}

// Warm-start transform for 64 < n <= 128: channel 0 of a point <- W^+ H W, H the Hermitian matrix of its upper
// triangle (the eigensolver's input), W the group anchor's eigenvectors (pass-1 points, uprev null) or the same
// point's of the previous mesh plane (uprev).  One 256-thread workgroup per point; the Hermitian matrix and
// W^+ H are staged in the point's solver scratch (its record and V regions, free before the solve).
template <int N>
__global__ __launch_bounds__(256) void k_eigb_warm(double *__restrict__ x, const double2 *__restrict__ u,
                                                   const double2 *__restrict__ uprev, int n, int ldo, int count,
                                                   double2 *__restrict__ scr) {
    using L = EigbLayout<N>;
    constexpr int nt = N / 16;
    int anchor = -1;
    const int k = uprev ? (int)blockIdx.x : wn_warm_k((int)blockIdx.x, 1, count, &anchor);
    if (k >= count) return;
    const double2 *w = uprev ? uprev + (size_t)k * n * n : u + (size_t)anchor * n * n;
    double2 *sk = scr + (size_t)k * L::PER_K;
    double2 *hs = sk + L::REC;   // Hermitian H [N][N]
    double2 *ts = sk;            // W^+ H [N][N]
    double2 *x0 = (double2 *)(x + (size_t)k * ldo);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lq = lane >> 4;
    for (int e = t; e < N * N; e += 256) {
        const int i = e / N, j = e - (e / N) * N;
        double2 v = make_double2(0.0, 0.0);
        if (i < n && j < n) {
            if (j >= i) {
                v = x0[i * n + j];
                if (i == j) v.y = 0.0;
            } else {
                v = x0[j * n + i];
                v.y = -v.y;
            }
        }
        hs[e] = v;
    }
    __threadfence_block();
    __syncthreads();
    // T = W^+ H: (W^+)(r, kk) = conj W(kk, r)
    for (int tile = wave; tile < nt * nt; tile += 4) {
        const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
        f64x4 p1, p2, p3;
        wn_mfma_tile_c64<N>(w + min(r0 + lr, n - 1), n, r0 + lr < n, n, -1.0, hs + c0 + lr, N, true, N, lq, p1, p2, p3);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            ts[(r0 + lq + 4 * q) * N + c0 + lr] = make_double2(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
    }
    __threadfence_block();
    __syncthreads();
    // channel 0 <- T W
    for (int tile = wave; tile < nt * nt; tile += 4) {
        const int r0 = (tile / nt) * 16, c0 = (tile % nt) * 16;
        f64x4 p1, p2, p3;
        wn_mfma_tile_c64<N>(ts + (r0 + lr) * N, 1, true, N, 1.0, w + min(c0 + lr, n - 1), n, c0 + lr < n, n, lq, p1, p2,
                            p3);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + lq + 4 * q, col = c0 + lr;
            if (row < n && col < n) x0[row * n + col] = make_double2(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
        }
    }
}

template <int N>
static hipError_t launch_eig_big(const double *x, double *e, double2 *u, int *sweeps,
                                 unsigned long long *sweep_total, int count, const WnShape &sh, hipStream_t st,
                                 double2 *scr, int ctr_stride, int *status, int pass, const double2 *uprev) {
    const double tol = 1.0e-15;
    const int n = sh.n;
    const int groups = (n + 4 * WN_EIGB_RW - 1) / (4 * WN_EIGB_RW);
    const int chunks = (sh.max_sweeps + WN_EIGB_SB - 1) / WN_EIGB_SB;
    const int nslots = eigb_slots(pass, count);
    const unsigned vgrid = (unsigned)((nslots + 7) / 8 * 8) * (unsigned)groups;
    for (int c = 0; c < chunks; ++c) {
        hipLaunchKernelGGL((k_eigb_sweeps<N, 512>), dim3(nslots), dim3(512), 0, st, x, n, sh.ldo, tol, count, scr, c,
                           sh.max_sweeps, pass);
        hipLaunchKernelGGL((k_eigb_vapply<N, WN_EIGB_RW>), dim3(vgrid), dim3(256), 0, st, n, count, scr, c, groups, pass,
                           nslots, u, uprev);
    }
    hipLaunchKernelGGL((k_eigb_finish<N>), dim3(nslots), dim3(256), 0, st, e, u, sweeps, sweep_total, n, count, scr,
                       ctr_stride, status, pass);
    return hipGetLastError();
}

hipError_t wn_launch_eig_big(const double *x, double *e, double2 *u, int *sweeps, unsigned long long *sweep_total,
                             int count, const WnShape &sh, hipStream_t st, double2 *scr, int ctr_stride, int *status,
                             int pass, const double2 *uprev) {
    if (sh.n <= 64 || sh.n > 128 || !scr || sh.max_sweeps < 1) return hipErrorInvalidValue;
    if ((pass == 0 || pass == 1 || pass == 2) && !wn_warm_start(sh.n, count)) return hipErrorInvalidValue;
    if (pass == 2 && !uprev) return hipErrorInvalidValue;
    if (count <= 0) return hipSuccess;
    if (sh.n <= 96)
        return launch_eig_big<96>(x, e, u, sweeps, sweep_total, count, sh, st, scr, ctr_stride, status, pass, uprev);
    return launch_eig_big<128>(x, e, u, sweeps, sweep_total, count, sh, st, scr, ctr_stride, status, pass, uprev);
}

hipError_t wn_launch_warm_big(double *x, const double2 *u, int count, const WnShape &sh, hipStream_t st,
                              const double2 *uprev, double2 *scr) {
    if (sh.n <= 64 || sh.n > 128 || !scr || !wn_warm_start(sh.n, count)) return hipErrorInvalidValue;
    const int grid = uprev ? count : (count + WN_WARM_G - 1) / WN_WARM_G * (WN_WARM_G - 1);
    if (sh.n <= 96)
        hipLaunchKernelGGL((k_eigb_warm<96>), dim3(grid), dim3(256), 0, st, x, u, uprev, sh.n, sh.ldo, count, scr);
    else
        hipLaunchKernelGGL((k_eigb_warm<128>), dim3(grid), dim3(256), 0, st, x, u, uprev, sh.n, sh.ldo, count, scr);
    return hipGetLastError();
}
