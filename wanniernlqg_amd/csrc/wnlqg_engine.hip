// Host side of libwnlqg_hip.so: the C ABI declared in include/wnlqg.h.
//
// One engine = one model + one resolved task on one device and one HIP stream.  wnlqg_create
// uploads the model and builds the R-space channels on the device once (cold, as
// prepare_real_space!, HamiltonianMatrixElements.jl:180-199); wnlqg_run_mesh streams the mesh in
// k batches through phase -> interpolation -> eigensolve -> rotation -> pairs -> contraction ->
// lane sums, all on the device, and returns only the lane tensors.

#include "../../include/wnlqg.h"
#include "wnlqg_common.h"
#include "wnlqg_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <vector>

#define WN_MARKS 11

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define WN_HIP(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess)                                                                             \
            return fail(_e == hipErrorOutOfMemory ? WNLQG_E_OOM : WNLQG_E_HIP, "%s failed: %s (%s:%d)", #expr, \
                        hipGetErrorString(_e), __FILE__, __LINE__);                                       \
    } while (0)

// SC / IC prefactors with kpoint_weight = num_kpoints = 1
// (ConventionalCurrentIntegralExecution.jl:243-245, :296-298; Core/Constants.jl:1-4).
const double kCharge = 1.602176634e-19;
const double kHbarJs = 1.054571817e-34;
const double kHbarEvs = 6.582119569e-16;
const double kPi = 3.14159265358979323846;

double sc_prefactor(double vol) {
    return kHbarEvs * kPi * (-kCharge) * (-kCharge) * (-kCharge) / (2.0 * kHbarJs * kHbarJs * 1 * vol);
}
double ic_prefactor(double vol) { return -kPi * kCharge * kCharge * kCharge / (kHbarJs * kHbarJs * 1 * vol); }

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// One batch of a mesh: k range [k0, k0 + cnt); separable: planes [k3, k3 + nk3) with k2 range
// [k2b, k2b + k2c) (several whole planes, or part of one plane).
struct MeshBatch {
    long long k0;
    int cnt, k3, k2b, k2c, nk3;
};

// Separable meshes group small k3 planes into batches of at least this many points, so that the
// one-workgroup-per-k kernels fill the GPU.
#define WN_MIN_MESH_BATCH 16384

// Device tables of one mesh integral, kept between calls with the same mesh and lane count.
struct MeshPlan {
    bool valid = false;
    int32_t mesh[3] = {0, 0, 0};
    int n_lanes = 0;
    long long nk = 0;
    bool sep = false;
    std::vector<MeshBatch> batches;
    std::vector<long long> bchunk, blane;
    DevBuf<int4> dchunks, dlanes;
};

}  // namespace

struct wnlqg_engine {
    int family = 0;     // 0 SC / IC, 1 BCD and/or intrinsic AHC (gp.kind), 3 SHG
    std::vector<double> elem_pref;  // per output element: the state's per-point prefactor
    std::vector<double> elem_norm;  // per output element: divisor of the total besides N_k (0: none)
    WnGeom gp{};        // BCD / AHC parameters
    WnShg sp{};         // SHG parameters
    DevBuf<double2> der, coef, one;  // SHG optical position derivatives, pair coefficients, one-band sums
    DevBuf<double2> escratch, rscratch;  // Jacobi records (n <= 32) or V (n > 64); rotation U^+ X (n > 64)
    DevBuf<int> pidx;                 // SHG pair index of every (n, m) (-1: inactive)
    DevBuf<double> pdiff;             // SHG pair energy differences
    int device = 0;
    hipStream_t stream = nullptr;
    WnShape sh{};
    WnScreen sc{};
    int precision = WNLQG_PRECISION_FP64;
    double cell_volume = 0.0;
    double pref[2] = {0.0, 0.0};
    int rmin[3] = {0, 0, 0};
    int batch = 0;      // k rows per batch (multiple of 128)
    int kpc = 16;       // k points per accumulation chunk

    DevBuf<double> hr, ar, rc, bf, kp, x, e, omega, wts;
    DevBuf<int> deg, sweeps, count, flags, ridxs;  // ridxs: k1 of a reduced mesh's representatives
    DevBuf<int4> ridx, pmeta, chunks, lanes, rchunks, rlanes, rgroups;
    DevBuf<double2> tab, u, pmd, q, pwd, partial, lanebuf, lanecomp;
    DevBuf<float2> tab32;
    DevBuf<__bf16> bh, bm, bl;  // split-BF16 coefficient parts (alternative mixed kernel)
    DevBuf<_Float16> fh, fm;    // split-FP16 coefficient parts (mixed forms)
    DevBuf<double> cup, cdown;  // per-column power-of-two scales of the FP16 operand
    DevBuf<float2> t12, t3;     // per-k (R1,R2)-plane and R3 phase tables (FP32)
    DevBuf<int2> ridx2;         // per R vector: (j1*S + j2, j3)
    int nterms = 0;             // 0 = FP64 interpolation, else split products per MFMA output
    int kind = 0;               // 0 FP64, 1 split FP16, 2 split BF16
    int backend = WNLQG_FOURIER_AUTO;
    bool separable = false;     // meshes use the separable plane contraction
    bool ch32 = false;          // mixed form, SC / IC on separable meshes: channels after H stored in FP32
    int span[3] = {1, 1, 1};    // R box extent per axis
    DevBuf<double> box;         // [S1*S2*S3][ldo] channel rows of the dense R box (FP64)
    DevBuf<double2> y3, y23, e1, e2, e3;
    int h_tiles = 0;            // 128-column FP64 tiles holding channel H (mixed forms)
    int partial_cap = 0;  // chunks the partial buffer holds
    int eig_tpb = 0;      // Jacobi threads per k point (0: default; set by the batch autotuner)
    // mesh pipeline: batch b is interpolated and diagonalized on `stream` into buffer b % 2 while
    // batch b - 1 is rotated and integrated on `stream2`
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_ready[2] = {}, ev_free[2] = {};
    MeshPlan plan;        // batches and chunk / lane tables of the last mesh
    std::vector<double> omega_host;
    hipEvent_t ev[8] = {};
};

namespace {

// driver family of a quantities bitmask: 0 SC / IC, 1 BCD and/or AHC, 3 SHG, -1 invalid
int task_family(int q) {
    const int geom = WNLQG_BERRY_CURVATURE_DIPOLE | WNLQG_INTRINSIC_HALL;
    if (q != 0 && (q & ~geom) == 0) return 1;
    if (q == WNLQG_SECOND_HARMONIC) return 3;
    if (q != 0 && (q & ~(WNLQG_SHIFT_CURRENT | WNLQG_INJECTION_CURRENT)) == 0) return 0;
    return -1;
}

// BCD / AHC task checks (BerryMultipoleDC.jl:16-55, IntrinsicHallOccupations.jl:16-21).
int check_geometry_task(const wnlqg_task *t) {
    const bool bcd = (t->quantities & WNLQG_BERRY_CURVATURE_DIPOLE) != 0;
    const bool ahc = (t->quantities & WNLQG_INTRINSIC_HALL) != 0;
    if (t->num_fermi_energies < 1 || t->num_fermi_energies > WN_GEOM_MAXMU || !t->fermi_energies)
        return fail(WNLQG_E_INVALID, "1 to %d Fermi energies are required", WN_GEOM_MAXMU);
    for (int i = 0; i < t->num_fermi_energies; ++i)
        if (!std::isfinite(t->fermi_energies[i])) return fail(WNLQG_E_NONFINITE, "non-finite Fermi energy");
    if (!(t->temperature >= 0.0) || !std::isfinite(t->temperature))
        return fail(WNLQG_E_INVALID, "temperature_k must be finite and nonnegative");
    if (!(t->denominator_regularization >= 0.0) || !std::isfinite(t->denominator_regularization))
        return fail(WNLQG_E_INVALID, "denominator_regularization must be finite and >= 0");
    if (t->precision != WNLQG_PRECISION_FP64)
        return fail(WNLQG_E_UNSUPPORTED, "BCD / AHC run in FP64 only");
    if (t->batch_size < 0) return fail(WNLQG_E_INVALID, "batch_size must be >= 0");
    if (t->fourier_backend < WNLQG_FOURIER_AUTO || t->fourier_backend > WNLQG_FOURIER_MIXED)
        return fail(WNLQG_E_INVALID, "unknown fourier_backend %d", t->fourier_backend);
    if (bcd) {
        if (!(t->degeneracy_threshold >= 0.0) || !std::isfinite(t->degeneracy_threshold))
            return fail(WNLQG_E_INVALID, "invalid spectral isolation threshold");
        const double w = t->fermi_surface_broadening;
        if (t->temperature == 0.0 && !(w > 0.0 && std::isfinite(w)))
            return fail(WNLQG_E_INVALID, "zero-temperature Fermi surface requires explicit positive eta_fs_ev");
        if (t->temperature > 0.0 && w != 0.0)
            return fail(WNLQG_E_INVALID, "eta_fs_ev is applicable only to zero-temperature Fermi surface");
    }
    if (ahc && t->temperature != 0.0) return fail(WNLQG_E_INVALID, "INTRINSIC_HALL_REQUIRES_ZERO_TEMPERATURE");
    return WNLQG_OK;
}

int check_task(const wnlqg_task *t, int n) {
    if (!t) return fail(WNLQG_E_INVALID, "task is NULL");
    const int family = task_family(t->quantities);
    if (family < 0)
        return fail(WNLQG_E_UNSUPPORTED,
                    "quantities bitmask %d: shift (1) and/or injection (2) current, or BCD (4) and/or AHC (8), or SHG (16)",
                    t->quantities);
    if (family == 1) return check_geometry_task(t);
    if (family == 3) {
        if (t->num_shg_components < 0 || t->num_shg_components > 18 || (t->num_shg_components > 0 && !t->shg_components))
            return fail(WNLQG_E_INVALID, "0 to 18 SHG components");
        for (int i = 0; i < 3 * t->num_shg_components; ++i)
            if (t->shg_components[i] < 0 || t->shg_components[i] > 2)
                return fail(WNLQG_E_INVALID, "SHG component indices must be 0, 1 or 2");
        if (!(t->shg_low_frequency_broadening > 0.0) || !std::isfinite(t->shg_low_frequency_broadening))
            return fail(WNLQG_E_INVALID, "shg_low_frequency_broadening must be finite and positive");
        if (!(t->shg_intermediate_regularization > 0.0) || !std::isfinite(t->shg_intermediate_regularization))
            return fail(WNLQG_E_INVALID, "shg_intermediate_regularization must be finite and positive");
        if (!(t->degeneracy_threshold >= 0.0) || !std::isfinite(t->degeneracy_threshold))
            return fail(WNLQG_E_INVALID, "invalid degeneracy threshold");
        if (t->precision != WNLQG_PRECISION_FP64) return fail(WNLQG_E_UNSUPPORTED, "SHG runs in FP64 only");
    }
    if (t->num_photon_energies < 1 || !t->photon_energies)
        return fail(WNLQG_E_INVALID, "photon energies missing");
    for (int i = 0; i < t->num_photon_energies; ++i) {
        if (!std::isfinite(t->photon_energies[i])) return fail(WNLQG_E_NONFINITE, "non-finite photon energy");
        if (i > 0 && t->photon_energies[i] < t->photon_energies[i - 1])
            return fail(WNLQG_E_UNSUPPORTED, "photon energies must be ascending");
    }
    if (t->broadening_type != WNLQG_GAUSSIAN && t->broadening_type != WNLQG_LORENTZIAN)
        return fail(WNLQG_E_INVALID, "Unknown smearing type: %d", t->broadening_type);
    if (!(t->broadening > 0.0) || !std::isfinite(t->broadening))
        return fail(WNLQG_E_INVALID, "broadening must be finite and positive");
    if (!(t->temperature >= 0.0) || !std::isfinite(t->temperature) || !std::isfinite(t->fermi_energy))
        return fail(WNLQG_E_INVALID, "invalid temperature or Fermi energy");
    if (!(t->denominator_regularization >= 0.0) || !std::isfinite(t->denominator_regularization))
        return fail(WNLQG_E_INVALID, "denominator_regularization must be finite and >= 0");
    if (t->band_window_size != -1 && 2 * t->band_window_size <= n)
        return fail(WNLQG_E_UNSUPPORTED, "band_window_size %d selects a partial window; only all bands are implemented",
                    t->band_window_size);
    if (t->precision != WNLQG_PRECISION_FP64 && t->precision != WNLQG_PRECISION_MIXED &&
        t->precision != WNLQG_PRECISION_MIXED_HIGH && t->precision != WNLQG_PRECISION_SPLIT_BF16)
        return fail(WNLQG_E_INVALID, "unknown precision %d", t->precision);
    if (t->batch_size < 0) return fail(WNLQG_E_INVALID, "batch_size must be >= 0");
    if (t->fourier_backend < WNLQG_FOURIER_AUTO || t->fourier_backend > WNLQG_FOURIER_MIXED)
        return fail(WNLQG_E_INVALID, "unknown fourier_backend %d", t->fourier_backend);
    return WNLQG_OK;
}

// complex values per k point (or chunk, or lane) of all output states
int output_elems(const wnlqg_engine *g) {
    if (g->family == 1) return ((g->gp.kind & 1) ? 27 : 0) * g->gp.nmu + ((g->gp.kind & 2) ? 9 : 0) * g->gp.nmu;
    if (g->family == 3) return 19 * g->sp.ncomp * g->sp.nw + g->sp.ncomp;
    return 27 * g->sc.nq * g->sc.nw;
}

// complex values per k point (or lane) handed to the caller
int result_elems(const wnlqg_engine *g) { return g->family == 3 ? 14 * g->sp.ncomp * g->sp.nw : output_elems(g); }

int ensure_rows(wnlqg_engine *g, int rows) {
    // working set for `rows` k points (multiple of 128)
    const size_t n2 = (size_t)g->sh.n * g->sh.n;
    if (g->x.n >= (size_t)2 * rows * g->sh.ldo) return WNLQG_OK;
    // channels, eigenvalues and eigenvectors are double-buffered for the two-stream mesh pipeline
    WN_HIP(g->x.alloc((size_t)2 * rows * g->sh.ldo));
    WN_HIP(g->kp.alloc((size_t)rows * 3));
    WN_HIP(g->tab.alloc((size_t)rows * 3 * g->sh.S));
    if (g->kind == 2) WN_HIP(g->tab32.alloc((size_t)rows * 3 * g->sh.S));
    if (g->kind == 1) {
        WN_HIP(g->t12.alloc((size_t)rows * g->sh.S * g->sh.S));
        WN_HIP(g->t3.alloc((size_t)rows * g->sh.S));
    }
    WN_HIP(g->e.alloc((size_t)2 * rows * g->sh.n));
    WN_HIP(g->u.alloc((size_t)2 * rows * n2));
    WN_HIP(g->sweeps.alloc((size_t)rows));
    if (wn_eig_scratch_elems(g->sh.n, rows)) WN_HIP(g->escratch.alloc(wn_eig_scratch_elems(g->sh.n, rows)));
    if (g->sh.n > 64) WN_HIP(g->rscratch.alloc(wn_rotate_scratch_elems(g->sh.n, rows)));
    if (g->family == 3) {
        WN_HIP(g->count.alloc((size_t)rows));
        WN_HIP(g->der.alloc((size_t)rows * WN_SHG_DER_BLOCKS * n2));   // 9 derivative blocks + 12 operand blocks
        WN_HIP(g->pidx.alloc((size_t)rows * n2));
        WN_HIP(g->pdiff.alloc((size_t)rows * n2));
        WN_HIP(g->coef.alloc((size_t)rows * n2 * 19 * g->sp.ncomp));
        WN_HIP(g->one.alloc((size_t)rows * g->sp.ncomp));
    }
    if (g->family == 0) {
        WN_HIP(g->count.alloc((size_t)rows));
        WN_HIP(g->pmeta.alloc((size_t)rows * n2));
        WN_HIP(g->pmd.alloc((size_t)rows * n2));
        WN_HIP(g->q.alloc((size_t)rows * n2 * 27 * g->sc.nq));
        WN_HIP(g->pwd.alloc((size_t)rows * WN_PCAP * 30));
    }
    return WNLQG_OK;
}

int ensure_partial(wnlqg_engine *g, int nchunks) {
    if (g->partial_cap >= nchunks) return WNLQG_OK;
    const size_t nelem = (size_t)output_elems(g);
    WN_HIP(g->partial.alloc((size_t)nchunks * nelem));
    WN_HIP(g->chunks.alloc((size_t)nchunks));
    g->partial_cap = nchunks;
    return WNLQG_OK;
}

hipError_t launch_phases(wnlqg_engine *g, int rows) {
    hipError_t e = wn_launch_phase_tables(g->kp.p, g->rmin, g->tab.p, g->kind == 2 ? g->tab32.p : nullptr, rows,
                                          g->sh.S, g->stream);
    if (e == hipSuccess && g->kind == 1) e = wn_launch_phase_planes(g->tab.p, g->t12.p, g->t3.p, rows, g->sh.S, g->stream);
    return e;
}

// Mixed forms: every channel after H in split BF16, then H (the eigenproblem input, columns
// [0, 2n^2)) in FP64, which also overwrites the BF16 values of a shared boundary tile.
hipError_t launch_interp(wnlqg_engine *g, int rows, double *out) {
    if (g->nterms) {
        const int hcols = 2 * g->sh.n * g->sh.n;
        hipError_t e = g->kind == 1
                           ? wn_launch_interp_f16(g->fh.p, g->fm.p, g->cdown.p, g->t12.p, g->t3.p, g->ridx2.p, out,
                                                  rows, g->nterms, hcols, g->sh, g->stream)
                           : wn_launch_interp_bf16(g->bh.p, g->bm.p, g->bl.p, g->tab32.p, g->ridx.p, out, rows,
                                                   g->nterms, hcols, g->sh, g->stream);
        if (e != hipSuccess) return e;
        WnShape hs = g->sh;
        hs.nNt = g->h_tiles;
        return wn_launch_interp_f64(g->bf.p, g->tab.p, g->ridx.p, out, rows, hs, g->stream);
    }
    return wn_launch_interp_f64(g->bf.p, g->tab.p, g->ridx.p, out, rows, g->sh, g->stream);
}

// Interpolation + eigensystem for `count` k points already in g->kp (rows padded to 128).
int run_front(wnlqg_engine *g, int count, int rows, bool eig) {
    WN_HIP(launch_phases(g, rows));
    WN_HIP(launch_interp(g, rows, g->x.p));
    if (eig)
        WN_HIP(wn_launch_eig_f64(g->x.p, g->e.p, g->u.p, g->sweeps.p, count, g->sh, g->eig_tpb, g->stream,
                                 g->escratch.p, g->flags.p));
    return WNLQG_OK;
}

int check_flags(wnlqg_engine *g);
hipError_t clear_flags(wnlqg_engine *g);

int upload_points(wnlqg_engine *g, const double *kpoints, int m, int rows) {
    std::vector<double> k((size_t)rows * 3, 0.0);
    for (int i = 0; i < m; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = kpoints[3 * i + a];
            if (!std::isfinite(v)) return fail(WNLQG_E_NONFINITE, "non-finite k point %d", i);
            k[3 * i + a] = v;
        }
    // the previous batch may still read kp: drain the stream, then copy synchronously
    WN_HIP(hipStreamSynchronize(g->stream));
    WN_HIP(hipMemcpy(g->kp.p, k.data(), k.size() * sizeof(double), hipMemcpyHostToDevice));
    return WNLQG_OK;
}

}  // namespace

extern "C" {

int32_t wnlqg_version(void) { return WNLQG_ABI_VERSION; }

int32_t wnlqg_last_error(char *buf, size_t len) {
    if (buf && len) {
        size_t c = std::min(len - 1, g_last_error.size());
        memcpy(buf, g_last_error.data(), c);
        buf[c] = 0;
    }
    return (int32_t)g_last_error.size();
}

int32_t wnlqg_num_states(const wnlqg_engine *g) { return g ? g->sc.nq : 0; }

namespace {

// Frequency-only factors of the SHG lane sums (SecondHarmonicKernels.jl:86-98, 200-236, 280-292):
// acc = [19 nc][nw] column sums then [nc] (one-band BCD, semiclassical) sums -> [nw, nc, 7, 2] column-major.
void shg_finalize(const wnlqg_engine *g, const double2 *acc, double *out) {
    const WnShg &sp = g->sp;
    const int nc = sp.ncomp, nw = sp.nw;
    const double2 *one = acc + (size_t)19 * nc * nw;
    for (size_t i = 0; i < (size_t)nw * nc * 14 * 2; ++i) out[i] = 0.0;
    auto at = [&](int w, int ci, int t, int q) { return out + 2 * ((size_t)w + (size_t)nw * (ci + (size_t)nc * (t + 7 * q))); };
    auto add = [&](int w, int ci, int t, int q, double re, double im) {
        double *o = at(w, ci, t, q);
        o[0] += re;
        o[1] += im;
    };
    for (int w = 0; w < nw; ++w) {
        const double om = g->omega_host[w];
        const double lb = sp.lowb;
        const double low = om / (om * om + lb * lb);
        const double y = om / lb;
        const double dl = sp.gaussian ? std::exp(-y * y) / (std::sqrt(kPi) * lb) : 1.0 / (kPi * lb * (1.0 + y * y));
        // kernel = low - i delta
        const double kr = low, ki = -dl;
        for (int ci = 0; ci < nc; ++ci) {
            auto col = [&](int c) { return acc[(size_t)c * nw + w]; };
            const int d1 = ci * 8, d2 = 8 * nc + ci * 6, dn = 14 * nc + ci * 2, lw = 16 * nc + ci * 3;
            double2 v;
            v = col(d1 + 0); add(w, ci, 0, 0, v.x, v.y);
            v = col(d1 + 1); add(w, ci, 1, 0, v.x, v.y);
            v = col(d1 + 2); add(w, ci, 3, 0, v.x, v.y);
            v = col(d1 + 3); add(w, ci, 4, 0, v.x * low * low, v.y * low * low);
            v = col(d1 + 4); add(w, ci, 0, 1, v.x, v.y);
            v = col(d1 + 5); add(w, ci, 1, 1, v.x, v.y);
            v = col(d1 + 6); add(w, ci, 3, 1, v.x, v.y);
            v = col(d1 + 7); add(w, ci, 4, 1, v.x * low, v.y * low);
            v = col(d2 + 0); add(w, ci, 0, 0, v.x, v.y);
            v = col(d2 + 1); add(w, ci, 1, 0, v.x, v.y);
            v = col(d2 + 2); add(w, ci, 2, 0, v.x, v.y);
            v = col(d2 + 3); add(w, ci, 0, 1, v.x, v.y);
            v = col(d2 + 4); add(w, ci, 1, 1, v.x, v.y);
            v = col(d2 + 5); add(w, ci, 2, 1, v.x, v.y);
            v = col(dn + 0); add(w, ci, 2, 0, v.x, v.y);
            v = col(dn + 1); add(w, ci, 2, 1, v.x, v.y);
            v = col(lw + 0); add(w, ci, 0, 0, v.x * low, v.y * low);
            v = col(lw + 1); add(w, ci, 1, 0, v.x * low, v.y * low);
            v = col(lw + 2); add(w, ci, 3, 0, v.x * low, v.y * low);
            const double bcd = one[ci].x, semi = one[ci].y;
            // t5q0 -= bcd kernel low / 4;  t5q1 += i bcd kernel / 2
            add(w, ci, 5, 0, -0.25 * bcd * kr * low, -0.25 * bcd * ki * low);
            add(w, ci, 5, 1, -0.5 * bcd * ki, 0.5 * bcd * kr);
            // t6q0 -= i semi kernel low^2 / 4;  t6q1 -= semi kernel low / 2
            add(w, ci, 6, 0, 0.25 * semi * ki * low * low, -0.25 * semi * kr * low * low);
            add(w, ci, 6, 1, -0.5 * semi * kr * low, -0.5 * semi * ki * low);
        }
    }
}

}  // namespace
int32_t wnlqg_num_channels(const wnlqg_engine *g) { return g ? g->sh.C : 0; }

int32_t wnlqg_destroy(wnlqg_engine *g) {
    if (!g) return WNLQG_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto &e : g->ev)
        if (e) (void)hipEventDestroy(e);
    g->hr.release(); g->ar.release(); g->rc.release(); g->bf.release(); g->kp.release(); g->x.release();
    g->e.release(); g->omega.release(); g->deg.release(); g->sweeps.release(); g->count.release();
    g->flags.release(); g->ridx.release(); g->pmeta.release(); g->chunks.release(); g->lanes.release();
    g->ridxs.release(); g->rchunks.release(); g->rlanes.release(); g->rgroups.release(); g->wts.release();
    g->tab.release(); g->u.release(); g->pmd.release(); g->q.release(); g->partial.release();
    g->lanebuf.release(); g->lanecomp.release(); g->tab32.release(); g->bh.release(); g->bm.release();
    g->bl.release(); g->fh.release(); g->fm.release(); g->cup.release(); g->cdown.release(); g->t12.release();
    g->t3.release(); g->ridx2.release(); g->box.release(); g->y3.release(); g->y23.release(); g->e1.release();
    g->e2.release(); g->e3.release(); g->plan.dchunks.release(); g->plan.dlanes.release();
    g->der.release(); g->coef.release(); g->one.release(); g->pidx.release(); g->pdiff.release();
    for (hipEvent_t e : {g->ev_fork, g->ev_join, g->ev_ready[0], g->ev_ready[1], g->ev_free[0], g->ev_free[1]})
        if (e) (void)hipEventDestroy(e);
    if (g->stream2) (void)hipStreamDestroy(g->stream2);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return WNLQG_OK;
}

// Exact Hermitian symmetry of a model: every R has its -R with the same degeneracy, H(-R) = H(R)^+ and
// A_a(-R) = A_a(R)^+ entry for entry (models built that way, such as the synthetic ones, satisfy it bitwise;
// a model that is Hermitian only to rounding is treated as general).
static bool model_hermitian(const wnlqg_model *m) {
    const int n = m->num_orbitals, nr = m->num_r_vectors;
    std::map<std::array<int, 3>, int> index;
    for (int r = 0; r < nr; ++r)
        index[{m->r_vectors[3 * r], m->r_vectors[3 * r + 1], m->r_vectors[3 * r + 2]}] = r;
    const size_t n2 = (size_t)n * n;
    for (int r = 0; r < nr; ++r) {
        auto it = index.find({-m->r_vectors[3 * r], -m->r_vectors[3 * r + 1], -m->r_vectors[3 * r + 2]});
        if (it == index.end()) return false;
        const int q = it->second;
        if (m->r_degeneracies[q] != m->r_degeneracies[r]) return false;
        // column-major [n, n, ...]: entry (i, j) at i + j n
        for (int blk = 0; blk < (m->position_r ? 4 : 1); ++blk) {
            const double *a = blk == 0 ? m->hamiltonian_r + 2 * n2 * r : m->position_r + 2 * n2 * (3 * (size_t)r + blk - 1);
            const double *b = blk == 0 ? m->hamiltonian_r + 2 * n2 * q : m->position_r + 2 * n2 * (3 * (size_t)q + blk - 1);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const size_t ij = 2 * (i + (size_t)j * n), ji = 2 * (j + (size_t)i * n);
                    if (a[ij] != b[ji] || a[ij + 1] != -b[ji + 1]) return false;
                }
        }
    }
    return true;
}

int32_t wnlqg_create(const wnlqg_model *model, const wnlqg_task *task, wnlqg_engine **out) {
    if (!out) return fail(WNLQG_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!model || !model->lattice || !model->r_vectors || !model->r_degeneracies || !model->hamiltonian_r)
        return fail(WNLQG_E_INVALID, "model pointers missing");
    const int n = model->num_orbitals, nr = model->num_r_vectors;
    if (n < 1 || nr < 1) return fail(WNLQG_E_INVALID, "num_orbitals and num_r_vectors must be positive");
    if (n > 128) return fail(WNLQG_E_UNSUPPORTED, "num_orbitals %d > 128 is not supported by this build", n);
    int rc = check_task(task, n);
    if (rc) return rc;
    if (n > 64 && task_family(task->quantities) != 3)
        return fail(WNLQG_E_UNSUPPORTED, "num_orbitals %d > 64 is supported for SHG only (SC/IC/BCD/AHC: <= 64)", n);
    const size_t n2 = (size_t)n * n;
    for (size_t i = 0; i < 2 * n2 * nr; ++i)
        if (!std::isfinite(model->hamiltonian_r[i])) return fail(WNLQG_E_NONFINITE, "non-finite H(R) entry %zu", i);
    if (model->position_r)
        for (size_t i = 0; i < 2 * n2 * 3 * nr; ++i)
            if (!std::isfinite(model->position_r[i])) return fail(WNLQG_E_NONFINITE, "non-finite A(R) entry %zu", i);
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(model->lattice[i])) return fail(WNLQG_E_NONFINITE, "non-finite lattice");
    for (int r = 0; r < nr; ++r)
        if (model->r_degeneracies[r] <= 0) return fail(WNLQG_E_INVALID, "R degeneracy %d must be positive", r);

    wnlqg_engine *g = new wnlqg_engine();
    auto bail = [&](int code) {
        wnlqg_destroy(g);
        return code;
    };
    if (hipGetDevice(&g->device) != hipSuccess) return bail(fail(WNLQG_E_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&g->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(WNLQG_E_HIP, "stream creation failed"));
    for (hipEvent_t *e : {&g->ev_fork, &g->ev_join, &g->ev_ready[0], &g->ev_ready[1], &g->ev_free[0], &g->ev_free[1]})
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(WNLQG_E_HIP, "event creation failed"));
    for (auto &e : g->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(WNLQG_E_HIP, "event creation failed"));

    // lattice: Julia column-major lattice[i, a] at i + 3a; rows are lattice vectors
    const double *L = model->lattice;
    g->cell_volume = L[0] * (L[4] * L[8] - L[7] * L[5]) - L[3] * (L[1] * L[8] - L[7] * L[2]) +
                     L[6] * (L[1] * L[5] - L[4] * L[2]);
    g->precision = task->precision;
    g->family = task_family(task->quantities);
    if (g->family == 1) {
        g->gp.kind = ((task->quantities & WNLQG_BERRY_CURVATURE_DIPOLE) ? 1 : 0) |
                     ((task->quantities & WNLQG_INTRINSIC_HALL) ? 2 : 0);
        g->gp.nmu = task->num_fermi_energies;
        g->gp.temperature = task->temperature;
        g->gp.eta_fs = task->fermi_surface_broadening;
        g->gp.eta = task->denominator_regularization;
        g->gp.gap_tol = task->degeneracy_threshold;
    }
    if (g->family == 3) {
        static const int kDefault[18][3] = {{0, 0, 0}, {0, 1, 1}, {0, 2, 2}, {0, 1, 2}, {0, 0, 2}, {0, 0, 1},
                                            {1, 0, 0}, {1, 1, 1}, {1, 2, 2}, {1, 1, 2}, {1, 0, 2}, {1, 0, 1},
                                            {2, 0, 0}, {2, 1, 1}, {2, 2, 2}, {2, 1, 2}, {2, 0, 2}, {2, 0, 1}};
        WnShg &sp = g->sp;
        sp.ncomp = task->num_shg_components ? task->num_shg_components : 18;
        for (int i = 0; i < sp.ncomp; ++i) {
            const int *c = task->num_shg_components ? task->shg_components + 3 * i : kDefault[i];
            sp.comp[i] = c[0] + 3 * c[1] + 9 * c[2];
        }
        sp.nw = task->num_photon_energies;
        sp.gaussian = task->broadening_type == WNLQG_GAUSSIAN;
        sp.width = task->broadening;
        sp.lowb = task->shg_low_frequency_broadening;
        sp.eta = task->shg_intermediate_regularization;
        sp.eta_correction = task->shg_eta_correction != 0;
        sp.mu = task->fermi_energy;
        sp.temperature = task->temperature;
        sp.tol = task->degeneracy_threshold;
        // about 8k active pairs (n^2 / 2 per point) per accumulation chunk, so that large n still gives
        // the contraction many more chunk tasks than the chip has SIMDs
        g->kpc = std::max(1, std::min(16, (int)std::lround(16384.0 / ((double)n * n))));
    }
#ifndef WN_KPC_SCIC
#define WN_KPC_SCIC 32   // SC/IC: k points per accumulation chunk (16 -> 32: 51.5 -> 41.0 ms per 128^3 mesh)
#endif
    if (g->family == 0) g->kpc = WN_KPC_SCIC;
    g->sc.nw = g->family == 1 ? task->num_fermi_energies : task->num_photon_energies;
    g->sc.gaussian = task->broadening_type == WNLQG_GAUSSIAN;
    g->sc.width = task->broadening;
    g->sc.mu = task->fermi_energy;
    g->sc.temperature = task->temperature;
    g->sc.eta = task->denominator_regularization;
    g->sc.window = task->transition_window_factor * task->broadening;
    g->sc.sc = (task->quantities & WNLQG_SHIFT_CURRENT) != 0;
    g->sc.ic = (task->quantities & WNLQG_INJECTION_CURRENT) != 0;
    g->sc.nq = g->family == 1 ? ((g->gp.kind & 1) + ((g->gp.kind >> 1) & 1)) : (g->family == 3 ? 1 : g->sc.sc + g->sc.ic);
    int qi = 0;
    if (g->sc.sc) g->pref[qi++] = sc_prefactor(g->cell_volume);
    if (g->sc.ic) g->pref[qi++] = ic_prefactor(g->cell_volume);
    // per-element prefactors and total divisors: SC / IC states [27 n_w] each (prefactor, 1/N_k on the lanes);
    // BCD [27 n_mu] raw moments, total / (N_k |det L|); AHC [9 n_mu] with the per-point scale
    // -e^2/hbar 1e-20 (SpectralResponseDriver.jl:388-400), total / (N_k |det L| 1e-30)
    if (g->family == 0) {
        for (int q = 0; q < g->sc.nq; ++q)
            for (int e = 0; e < 27 * g->sc.nw; ++e) {
                g->elem_pref.push_back(g->pref[q]);
                g->elem_norm.push_back(0.0);
            }
    } else if (g->family == 1) {
        const double vol = std::fabs(g->cell_volume);
        if (g->gp.kind & 1)
            for (int e = 0; e < 27 * g->gp.nmu; ++e) {
                g->elem_pref.push_back(1.0);
                g->elem_norm.push_back(vol);
            }
        if (g->gp.kind & 2)
            for (int e = 0; e < 9 * g->gp.nmu; ++e) {
                g->elem_pref.push_back(-kCharge * kCharge / kHbarJs * 1e-20);
                g->elem_norm.push_back(vol * 1.0e-30);
            }
    }
    // the frequency axis of SC / IC / SHG, or the Fermi-energy axis of BCD / AHC
    const double *axis = g->family == 1 ? task->fermi_energies : task->photon_energies;
    g->omega_host.assign(axis, axis + g->sc.nw);

    // shape
    g->sh.n = n;
    g->sh.nr = nr;
    g->sh.has_pos = model->position_r != nullptr;
    // stored channels: SC / IC need H, dH, ddH (+ A, dA); BCD / AHC need H, dH (+ A, dA)
    {
        int c = 0;
        for (int l = 0; l < WN_CH_FULL; ++l) {
            const bool pos = l >= WN_CH_A;
            const bool second = l >= WN_CH_DDH && l < WN_CH_A;
            if (pos && !g->sh.has_pos) continue;
            if (second && g->family == 1) continue;
            g->sh.lch[c++] = (signed char)l;
        }
        g->sh.C = c;
        g->gp.c_dh = 1;
        g->gp.c_a = g->sh.has_pos ? (g->family == 1 ? 4 : WN_CH_A) : -1;
        g->gp.c_da = g->sh.has_pos ? (g->family == 1 ? 7 : WN_CH_DA) : -1;
    }
    g->sh.Kb = (nr + WN_RB - 1) / WN_RB;
    // columns padded to 256 so that both interpolation kernels tile the same output rows
    g->sh.nNt = (int)((2 * (size_t)g->sh.C * n2 + 255) / 256) * 2;
    // mixed forms: split FP16 with 3 (MIXED) or 4 (MIXED_HIGH) products, split BF16 with 3 (SPLIT_BF16)
    g->nterms = task->precision == WNLQG_PRECISION_MIXED || task->precision == WNLQG_PRECISION_SPLIT_BF16
                    ? 3
                    : (task->precision == WNLQG_PRECISION_MIXED_HIGH ? 4 : 0);
    g->kind = task->precision == WNLQG_PRECISION_SPLIT_BF16 ? 2 : (g->nterms ? 1 : 0);
    g->h_tiles = (int)((2 * n2 + WN_GN - 1) / WN_GN);
    g->sh.ldo = g->sh.nNt * WN_GN;
    int rmax[3];
    for (int a = 0; a < 3; ++a) {
        g->rmin[a] = model->r_vectors[a];
        rmax[a] = model->r_vectors[a];
    }
    for (int r = 0; r < nr; ++r)
        for (int a = 0; a < 3; ++a) {
            g->rmin[a] = std::min(g->rmin[a], model->r_vectors[3 * r + a]);
            rmax[a] = std::max(rmax[a], model->r_vectors[3 * r + a]);
        }
    // Jacobi sweep cap (a point that reaches it fails the call with WNLQG_E_NOT_CONVERGED); converged
    // solves take 4-8 sweeps.  WNLQG_JACOBI_MAX_SWEEPS lowers it (tests force the failure path with it).
    g->sh.max_sweeps = 40;
    if (const char *ms = std::getenv("WNLQG_JACOBI_MAX_SWEEPS")) {
        const int v = std::atoi(ms);
        if (v >= 1 && v <= 40) g->sh.max_sweeps = v;
    }
    // deferred-eigenvector Jacobi: sweeps recorded per point (1..8) before a point is finished by the register-V
    // kernel from its saved state; WNLQG_EIGD_RECORD_SWEEPS lowers it (tests drive the continuation with it)
    g->sh.rec_sweeps = 8;
    if (const char *rs = std::getenv("WNLQG_EIGD_RECORD_SWEEPS")) {
        const int v = std::atoi(rs);
        if (v >= 1 && v <= 8) g->sh.rec_sweeps = v;
    }
    // rows-only element rotation for exactly Hermitian models (WNLQG_ROTATE_HERMITIAN=0 turns it off)
    g->sh.herm = model_hermitian(model) ? 1 : 0;
    if (const char *hv = std::getenv("WNLQG_ROTATE_HERMITIAN"))
        if (!strcmp(hv, "0")) g->sh.herm = 0;
    // eigensolver form: WNLQG_EIG_KERNEL = "lds64" / "lds128" / "lds256" (LDS kernel with that many threads per
    // point), "regv" (register-resident eigenvectors, n_w <= 32) or "defer" (H-only sweeps with recorded rotations
    // applied to V afterwards, n_w <= 32); default: the size class's default ("defer" for n_w <= 32)
    if (const char *ek = std::getenv("WNLQG_EIG_KERNEL")) {
        if (!strcmp(ek, "lds64")) g->eig_tpb = 64;
        else if (!strcmp(ek, "lds128")) g->eig_tpb = 128;
        else if (!strcmp(ek, "lds256")) g->eig_tpb = 256;
        else if (!strcmp(ek, "regv")) g->eig_tpb = WN_EIG_RV;
        else if (!strcmp(ek, "defer")) g->eig_tpb = WN_EIG_DV;
    }
    g->sh.S = 1;
    for (int a = 0; a < 3; ++a) {
        g->span[a] = rmax[a] - g->rmin[a] + 1;
        g->sh.S = std::max(g->sh.S, g->span[a]);
    }
    g->backend = task->fourier_backend;

    // host-side R tables: Cartesian R (as _r_vector_cartesian, HamiltonianMatrixElements.jl:74-85)
    std::vector<double> rcart((size_t)3 * nr);
    std::vector<int4> rid((size_t)g->sh.Kb * WN_RB, make_int4(0, 0, 0, 0));
    std::vector<int2> rid2((size_t)g->sh.Kb * WN_RB, make_int2(0, 0));
    for (int r = 0; r < nr; ++r) {
        const int *R = model->r_vectors + 3 * r;
        for (int a = 0; a < 3; ++a)
            rcart[3 * r + a] = R[0] * L[0 + 3 * a] + R[1] * L[1 + 3 * a] + R[2] * L[2 + 3 * a];
        rid[r] = make_int4(R[0] - g->rmin[0], R[1] - g->rmin[1], R[2] - g->rmin[2], 0);
        rid2[r] = make_int2((R[0] - g->rmin[0]) * g->sh.S + (R[1] - g->rmin[1]), R[2] - g->rmin[2]);
    }
    if (g->hr.alloc(2 * n2 * nr) || g->rc.alloc(3 * (size_t)nr) || g->deg.alloc(nr) ||
        g->ridx.alloc(rid.size()) || g->omega.alloc(g->sc.nw) || g->flags.alloc(WN_FLAG_WORDS) ||
        g->bf.alloc((size_t)g->sh.Kb * (g->nterms ? g->h_tiles : g->sh.nNt) * 2048) ||
        (g->kind == 2 && (g->bh.alloc((size_t)g->sh.Kb * 16 * g->sh.ldo) || g->bm.alloc((size_t)g->sh.Kb * 16 * g->sh.ldo))) ||
        (g->kind == 2 && g->nterms >= 4 && g->bl.alloc((size_t)g->sh.Kb * 16 * g->sh.ldo)) ||
        (g->kind == 1 && (g->fh.alloc((size_t)g->sh.Kb * 16 * g->sh.ldo) || g->fm.alloc((size_t)g->sh.Kb * 16 * g->sh.ldo) ||
                          g->cup.alloc(g->sh.ldo) || g->cdown.alloc(g->sh.ldo) || g->ridx2.alloc(rid.size()))) ||
        (g->sh.has_pos && g->ar.alloc(2 * n2 * 3 * nr)))
        return bail(fail(WNLQG_E_OOM, "device allocation for the model failed"));
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(g->hr.p, model->hamiltonian_r, 2 * n2 * nr * sizeof(double)) ||
        !up(g->rc.p, rcart.data(), rcart.size() * sizeof(double)) ||
        !up(g->deg.p, model->r_degeneracies, nr * sizeof(int)) ||
        !up(g->ridx.p, rid.data(), rid.size() * sizeof(int4)) ||
        (g->kind == 1 && !up(g->ridx2.p, rid2.data(), rid2.size() * sizeof(int2))) ||
        !up(g->omega.p, g->omega_host.data(), g->sc.nw * sizeof(double)) ||
        (g->sh.has_pos && !up(g->ar.p, model->position_r, 2 * n2 * 3 * nr * sizeof(double))))
        return bail(fail(WNLQG_E_HIP, "model upload failed"));
    hipError_t be = hipSuccess;
    if (g->nterms) {
        be = g->kind == 1 ? wn_launch_build_coeff_f16(g->hr.p, g->ar.p, g->rc.p, g->deg.p, g->cup.p, g->cdown.p,
                                                      g->fh.p, g->fm.p, g->sh, g->stream)
                          : wn_launch_build_coeff_bf16(g->hr.p, g->ar.p, g->rc.p, g->deg.p, g->bh.p, g->bm.p, g->bl.p,
                                                       g->sh, g->stream);
        WnShape hs = g->sh;
        hs.nNt = g->h_tiles;
        if (be == hipSuccess) be = wn_launch_build_coeff(g->hr.p, g->ar.p, g->rc.p, g->deg.p, g->bf.p, hs, g->stream);
    } else {
        be = wn_launch_build_coeff(g->hr.p, g->ar.p, g->rc.p, g->deg.p, g->bf.p, g->sh, g->stream);
    }
    if (be != hipSuccess) return bail(fail(WNLQG_E_HIP, "coefficient build launch failed"));
    if (hipStreamSynchronize(g->stream) != hipSuccess) return bail(fail(WNLQG_E_HIP, "coefficient build failed"));

    // dense R box for the separable mesh contraction (only if every box cell holds at most one R vector
    // and the spans fit the contraction kernel)
    if (g->backend != WNLQG_FOURIER_DIRECT) {
        const long long nbox = (long long)g->span[0] * g->span[1] * g->span[2];
        std::vector<int> boxmap((size_t)nbox, -1);
        bool ok = g->span[0] <= 16 && g->span[1] <= 16 && g->span[2] <= 16 && nbox <= 8 * (long long)nr + 4096;
        for (int r = 0; ok && r < nr; ++r) {
            const int *R = model->r_vectors + 3 * r;
            const long long b = ((long long)(R[0] - g->rmin[0]) * g->span[1] + (R[1] - g->rmin[1])) * g->span[2] +
                                (R[2] - g->rmin[2]);
            if (boxmap[b] >= 0) ok = false;
            boxmap[b] = r;
        }
        if (!ok && (g->backend == WNLQG_FOURIER_PLANES || g->backend == WNLQG_FOURIER_MIXED))
            return bail(fail(WNLQG_E_UNSUPPORTED, "separable contraction needs distinct R vectors in a box of span <= 16"));
        if (ok) {
            DevBuf<int> dmap;
            if (dmap.alloc(boxmap.size()) != hipSuccess || g->box.alloc((size_t)nbox * g->sh.ldo) != hipSuccess ||
                g->y3.alloc((size_t)g->span[0] * g->span[1] * (g->sh.ldo / 2)) != hipSuccess)
                return bail(fail(WNLQG_E_OOM, "device allocation for the R box failed"));
            if (!up(dmap.p, boxmap.data(), boxmap.size() * sizeof(int)) ||
                wn_launch_build_box(g->hr.p, g->ar.p, g->rc.p, g->deg.p, dmap.p, g->box.p, nbox, g->sh, g->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(g->stream) != hipSuccess)
                return bail(fail(WNLQG_E_HIP, "R box build failed"));
            g->separable = true;
            // the mixed form keeps the interpolated channels after H in FP32 on its mesh path (SC / IC, where the
            // element rotation reads them and writes back FP64)
            g->ch32 = g->family == 0 && g->precision == WNLQG_PRECISION_MIXED && wn_rotate_sc_elements(g->sh) &&
                      n % 2 == 0;
        }
    }

    // batch size: interpolation rows + pair buffers within ~24 GB unless the caller fixed it
    const size_t per_k = (size_t)g->sh.ldo * 8 +
                         (g->family == 0 ? n2 * (sizeof(int4) + sizeof(double2) + 27 * g->sc.nq * 16) : 0) +
                         (g->family == 3 ? n2 * (WN_SHG_DER_BLOCKS * 16 + 8 + 8 + 19 * (size_t)g->sp.ncomp * 16) : 0) +
                         n2 * 16 + (size_t)3 * g->sh.S * 16 +
                         (wn_eig_scratch_elems(n, 1) + wn_rotate_scratch_elems(n, 1)) * 16;
    long long b = task->batch_size > 0 ? task->batch_size : (long long)(24.0e9 / (double)per_k);
    b = std::max(128LL, std::min(b, 65536LL));
    g->batch = (int)((b + WN_GM - 1) / WN_GM * WN_GM);
    rc = ensure_rows(g, g->batch);
    if (rc) return bail(rc);
    *out = g;
    return WNLQG_OK;
}

int32_t wnlqg_synchronize(wnlqg_engine *g) {
    if (!g) return fail(WNLQG_E_INVALID, "engine is NULL");
    WN_HIP(hipStreamSynchronize(g->stream));
    return WNLQG_OK;
}

int32_t wnlqg_fourier_points(wnlqg_engine *g, const double *kpoints, int32_t m, double *out) {
    if (!g || !kpoints || !out || m < 0) return fail(WNLQG_E_INVALID, "invalid arguments");
    WN_HIP(hipSetDevice(g->device));
    const size_t blk = (size_t)g->sh.C * g->sh.n * g->sh.n * 2;
    for (int p0 = 0; p0 < m; p0 += g->batch) {
        int cnt = std::min(g->batch, m - p0);
        int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        int rc = upload_points(g, kpoints + 3 * (size_t)p0, cnt, rows);
        if (rc) return rc;
        rc = run_front(g, cnt, rows, false);
        if (rc) return rc;
        WN_HIP(hipMemcpy2DAsync(out + (size_t)p0 * blk, blk * sizeof(double), g->x.p, g->sh.ldo * sizeof(double),
                                blk * sizeof(double), cnt, hipMemcpyDeviceToHost, g->stream));
        WN_HIP(hipStreamSynchronize(g->stream));
    }
    return WNLQG_OK;
}

int32_t wnlqg_spectrum_points(wnlqg_engine *g, const double *kpoints, int32_t m, double *energies, double *vectors) {
    if (!g || !kpoints || !energies || !vectors || m < 0) return fail(WNLQG_E_INVALID, "invalid arguments");
    WN_HIP(hipSetDevice(g->device));
    WN_HIP(clear_flags(g));
    const size_t n = g->sh.n;
    for (int p0 = 0; p0 < m; p0 += g->batch) {
        int cnt = std::min(g->batch, m - p0);
        int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        int rc = upload_points(g, kpoints + 3 * (size_t)p0, cnt, rows);
        if (rc) return rc;
        rc = run_front(g, cnt, rows, true);
        if (rc) return rc;
        WN_HIP(hipMemcpyAsync(energies + (size_t)p0 * n, g->e.p, cnt * n * sizeof(double), hipMemcpyDeviceToHost,
                              g->stream));
        WN_HIP(hipMemcpyAsync(vectors + (size_t)p0 * n * n * 2, g->u.p, cnt * n * n * sizeof(double2),
                              hipMemcpyDeviceToHost, g->stream));
        WN_HIP(hipStreamSynchronize(g->stream));
    }
    return check_flags(g);
}

namespace {

int check_flags(wnlqg_engine *g) {
    int f[4] = {0, 0, 0, 0};
    WN_HIP(hipMemcpy(f, g->flags.p, sizeof(f), hipMemcpyDeviceToHost));
    if (f[1] & 1)
        return fail(WNLQG_E_NOT_CONVERGED, "Jacobi eigensolver reached %d sweeps without converging at some k point",
                    g->sh.max_sweeps);
    if (f[0] & 1)
        return fail(WNLQG_E_GAP,
                    "GEOMETRY_SINGLE_BAND_SPECTRUM_NOT_ISOLATED: denominator_regularization = 0 needs gaps > 1e-10 eV");
    if (f[2] & 1)
        return fail(WNLQG_E_GAP,
                    "GEOMETRY_SINGLE_BAND_SPECTRUM_NOT_ISOLATED: two levels within threshold_ev=%g at k index %d",
                    g->gp.gap_tol, f[3]);
    if (f[2] & 2) return fail(WNLQG_E_NONFINITE, "nonfinite BCD/BCQ or intrinsic Hall integrand");
    return WNLQG_OK;
}

// Hamiltonian gauge of cnt points in (x, e, u).  SC / IC: screening first, then the rotation of what the
// pairs use, then the pair kernels (the frequency contraction follows in launch_response); other
// families: every channel.
hipError_t launch_rotation(wnlqg_engine *g, double *x, const double *e, const double2 *u, int cnt, hipStream_t st) {
    if (g->family != 0) return wn_launch_rotate_f64(x, u, cnt, g->sh, st, g->rscratch.p);
    hipError_t r = wn_launch_screen(e, g->omega.p, g->count.p, g->pmeta.p, g->pmd.p, g->flags.p, cnt, g->sh, g->sc, st);
    if (r != hipSuccess) return r;
    r = wn_launch_rotate_sc(x, u, g->count.p, g->pmeta.p, g->pwd.p, cnt, g->sh, st, g->rscratch.p, g->flags.p);
    if (r != hipSuccess) return r;
    return wn_launch_pairs_f64(x, e, g->count.p, g->pmeta.p, wn_rotate_sc_elements(g->sh) ? g->pwd.p : nullptr,
                               g->q.p, cnt, g->sh, g->sc, st);
}

// clear the status words: [0] pair-gap flag, [2] geometry flags, [3] first failing k, and the counter
// buckets (max and total Jacobi sweeps, total active SC/IC pairs)
hipError_t clear_flags(wnlqg_engine *g) { return wn_launch_reset_flags(g->flags.p, g->stream); }

// per-chunk stage after the rotation: SC / IC pairs + frequency contraction, or BCD / AHC
hipError_t launch_response(wnlqg_engine *g, const int4 *chunks, int nch, int cnt, long long kbase) {
    if (g->family == 0)
        return wn_launch_accum(chunks, nch, g->count.p, g->pmeta.p, g->pmd.p, g->q.p, g->omega.p, g->partial.p, g->sh,
                               g->sc, g->stream);
    return wn_launch_geom_f64(g->x.p, g->e.p, g->omega.p, chunks, nch, nullptr, g->partial.p, g->flags.p, kbase, g->sh,
                              g->gp, g->stream);
}

}  // namespace

int32_t wnlqg_eval_points(wnlqg_engine *g, const double *kpoints, int32_t m, double *out) {
    return wnlqg_eval_points_spectrum(g, kpoints, m, out, nullptr);
}

int32_t wnlqg_eval_points_spectrum(wnlqg_engine *g, const double *kpoints, int32_t m, double *out,
                                   double *energies) {
    if (!g || !kpoints || !out || m < 0) return fail(WNLQG_E_INVALID, "invalid arguments");
    WN_HIP(hipSetDevice(g->device));
    const int nelem = output_elems(g);
    WN_HIP(clear_flags(g));
    std::vector<double> part;
    for (int p0 = 0; p0 < m; p0 += g->batch) {
        int cnt = std::min(g->batch, m - p0);
        int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        int rc = upload_points(g, kpoints + 3 * (size_t)p0, cnt, rows);
        if (rc) return rc;
        rc = run_front(g, cnt, rows, true);
        if (rc) return rc;
        if (energies)
            WN_HIP(hipMemcpyAsync(energies + (size_t)p0 * g->sh.n, g->e.p, (size_t)cnt * g->sh.n * sizeof(double),
                                  hipMemcpyDeviceToHost, g->stream));
        WN_HIP(launch_rotation(g, g->x.p, g->e.p, g->u.p, cnt, g->stream));
        rc = ensure_partial(g, cnt);
        if (rc) return rc;
        std::vector<int4> ch(cnt);
        for (int i = 0; i < cnt; ++i) ch[i] = make_int4(i, 1, i, 0);
        WN_HIP(hipStreamSynchronize(g->stream));
        WN_HIP(hipMemcpy(g->chunks.p, ch.data(), cnt * sizeof(int4), hipMemcpyHostToDevice));
        if (g->family == 3) {
            WN_HIP(wn_launch_shg(g->x.p, g->e.p, g->omega.p, g->chunks.p, cnt, cnt, g->der.p, g->count.p, g->pidx.p,
                                 g->pdiff.p, g->coef.p, g->one.p, g->partial.p, g->sh, g->sp, g->stream));
        } else {
            WN_HIP(launch_response(g, g->chunks.p, cnt, cnt, p0));
        }
        part.resize((size_t)cnt * nelem * 2);
        WN_HIP(hipMemcpyAsync(part.data(), g->partial.p, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              g->stream));
        WN_HIP(hipStreamSynchronize(g->stream));
        if (g->family == 3) {
            const size_t rel = (size_t)result_elems(g);
            for (int i = 0; i < cnt; ++i)
                shg_finalize(g, (const double2 *)(part.data() + (size_t)i * nelem * 2), out + (size_t)(p0 + i) * rel * 2);
            continue;
        }
        for (int i = 0; i < cnt; ++i)
            for (int e = 0; e < nelem; ++e) {
                const size_t src = ((size_t)i * nelem + e) * 2, dst = ((size_t)(p0 + i) * nelem + e) * 2;
                out[dst] = g->elem_pref[e] * part[src];
                out[dst + 1] = g->elem_pref[e] * part[src + 1];
            }
    }
    int rc = check_flags(g);
    if (rc) return rc;
    for (size_t i = 0; i < (size_t)m * result_elems(g) * 2; ++i)
        if (!std::isfinite(out[i])) return fail(WNLQG_E_NONFINITE, "non-finite response value");
    return WNLQG_OK;
}

}  // extern "C"

namespace {

// Batches, lane-aligned chunks and the device tables of a mesh (host work and allocations; not
// captured).  Lanes are contiguous k ranges (BundleAssembly.jl:41-47); each batch splits the part of
// every lane it covers into chunks of at most kpc points, in k order.
int plan_mesh(wnlqg_engine *g, const int32_t mesh[3], int32_t n_lanes) {
    MeshPlan &p = g->plan;
    if (p.valid && p.mesh[0] == mesh[0] && p.mesh[1] == mesh[1] && p.mesh[2] == mesh[2] && p.n_lanes == n_lanes)
        return WNLQG_OK;
    p.valid = false;
    if (g->kpc < 1 || g->kpc > WN_ACC_KMAX) return fail(WNLQG_E_INVALID, "chunk size %d outside [1, %d]", g->kpc, WN_ACC_KMAX);
    const long long nk = (long long)mesh[0] * mesh[1] * mesh[2];
    const int nelem = output_elems(g);
    std::vector<long long> lstart(n_lanes + 1);
    for (int l = 0; l <= n_lanes; ++l) lstart[l] = (long long)((__int128)l * nk / n_lanes);
    // batches: contiguous k ranges.  Separable: k2 ranges of one k3 plane (rows k2*N1 + k1).
    p.batches.clear();
    p.sep = g->separable && mesh[0] <= g->batch;
    if (!p.sep && g->backend == WNLQG_FOURIER_MIXED)
        return fail(WNLQG_E_UNSUPPORTED, "fourier_backend mixed: the separable contraction needs mesh axis 0 (%d) "
                    "within the batch (%d)", mesh[0], g->batch);
    if (p.sep) {
        const int k2step = std::max(1, std::min(mesh[1], g->batch / mesh[0]));
        const long long plane = (long long)mesh[0] * mesh[1];
        if (k2step == mesh[1]) {
            const int pp = (int)std::max<long long>(1, std::min<long long>(g->batch / plane,
                                                                           (WN_MIN_MESH_BATCH + plane - 1) / plane));
            for (int k3 = 0; k3 < mesh[2]; k3 += pp) {
                const int nk3 = std::min(pp, mesh[2] - k3);
                p.batches.push_back({(long long)k3 * plane, (int)(nk3 * plane), k3, 0, mesh[1], nk3});
            }
        } else {
            for (int k3 = 0; k3 < mesh[2]; ++k3)
                for (int k2b = 0; k2b < mesh[1]; k2b += k2step) {
                    const int k2c = std::min(k2step, mesh[1] - k2b);
                    p.batches.push_back({((long long)k3 * mesh[1] + k2b) * mesh[0], k2c * mesh[0], k3, k2b, k2c, 1});
                }
        }
        const size_t m2 = g->sh.ldo / 2;
        if (g->y23.n < (size_t)g->span[0] * k2step * m2) WN_HIP(g->y23.alloc((size_t)g->span[0] * k2step * m2));
        WN_HIP(g->e1.alloc((size_t)mesh[0] * g->span[0]));
        WN_HIP(g->e2.alloc((size_t)mesh[1] * g->span[1]));
        WN_HIP(g->e3.alloc((size_t)mesh[2] * g->span[2]));
        WN_HIP(wn_launch_axis_phases(g->e1.p, mesh[0], g->span[0], g->rmin[0], g->stream));
        WN_HIP(wn_launch_axis_phases(g->e2.p, mesh[1], g->span[1], g->rmin[1], g->stream));
        WN_HIP(wn_launch_axis_phases(g->e3.p, mesh[2], g->span[2], g->rmin[2], g->stream));
    } else {
        for (long long k0 = 0; k0 < nk; k0 += g->batch)
            p.batches.push_back({k0, (int)std::min<long long>(g->batch, nk - k0), 0, 0, 0, 0});
    }
    const long long nbatch = (long long)p.batches.size();
    std::vector<int4> all_chunks, all_lanes;
    p.bchunk.assign(nbatch + 1, 0);
    p.blane.assign(nbatch + 1, 0);
    int max_chunks = 0;
    for (long long bi = 0; bi < nbatch; ++bi) {
        const long long k0 = p.batches[bi].k0, k1 = k0 + p.batches[bi].cnt;
        p.bchunk[bi] = all_chunks.size();
        p.blane[bi] = all_lanes.size();
        int l = (int)(std::upper_bound(lstart.begin(), lstart.end(), k0) - lstart.begin()) - 1;
        for (; l < n_lanes && lstart[l] < k1; ++l) {
            long long a = std::max(k0, lstart[l]), b = std::min(k1, lstart[l + 1]);
            if (a >= b) continue;
            int cb = (int)(all_chunks.size() - p.bchunk[bi]);
            for (long long c = a; c < b; c += g->kpc)
                all_chunks.push_back(make_int4((int)(c - k0), (int)std::min<long long>(g->kpc, b - c), l, 0));
            int ce = (int)(all_chunks.size() - p.bchunk[bi]);
            all_lanes.push_back(make_int4(l, cb, ce, 0));
        }
        max_chunks = std::max(max_chunks, (int)(all_chunks.size() - p.bchunk[bi]));
    }
    p.bchunk[nbatch] = all_chunks.size();
    p.blane[nbatch] = all_lanes.size();
    int rc = ensure_partial(g, max_chunks);
    if (rc) return rc;
    WN_HIP(p.dchunks.alloc(all_chunks.size()));
    WN_HIP(p.dlanes.alloc(all_lanes.size()));
    WN_HIP(hipStreamSynchronize(g->stream));
    WN_HIP(hipMemcpy(p.dchunks.p, all_chunks.data(), all_chunks.size() * sizeof(int4), hipMemcpyHostToDevice));
    WN_HIP(hipMemcpy(p.dlanes.p, all_lanes.data(), all_lanes.size() * sizeof(int4), hipMemcpyHostToDevice));
    if (g->lanebuf.n < (size_t)n_lanes * nelem) {
        WN_HIP(g->lanebuf.alloc((size_t)n_lanes * nelem));
        WN_HIP(g->lanecomp.alloc((size_t)n_lanes * nelem));
    }
    for (int a = 0; a < 3; ++a) p.mesh[a] = mesh[a];
    p.n_lanes = n_lanes;
    p.nk = nk;
    p.valid = true;
    return WNLQG_OK;
}

// Eigensolver start of batch bi: 0 every point from the identity; 1 groups of 4 along the fastest
// axis (anchors from the identity, the rest from their anchor); 2 every point from the same point of
// the previous k3 plane (whole planes per batch, the previous plane in the other buffer).
int eig_scheme(const wnlqg_engine *g, long long bi) {
    const MeshPlan &p = g->plan;
    const MeshBatch &bt = p.batches[bi];
    if (!wn_warm_start(g->sh.n, bt.cnt)) return 0;
    if (p.sep && bi > 0 && bt.nk3 == 1 && p.batches[bi - 1].nk3 == 1 && bt.k2c == p.mesh[1] &&
        p.batches[bi - 1].k2c == p.mesh[1] && p.batches[bi - 1].cnt == bt.cnt && p.batches[bi - 1].k3 + 1 == bt.k3)
        return 2;
    return 1;
}

// Separable plane contraction of one batch into xb: plane k3 + j of the batch, k2 range [k2b, k2b + k2c),
// lands at rows j k2c N1 + (k2 - k2b) N1 + k1.
hipError_t launch_contract_batch(wnlqg_engine *g, const MeshBatch &bt, double *xb, hipStream_t s0) {
    const MeshPlan &p = g->plan;
    const long long m2 = g->sh.ldo / 2;
    const double2 *boxc = (const double2 *)g->box.p;
    hipError_t e = hipSuccess;
    for (int j = 0; j < bt.nk3 && e == hipSuccess; ++j) {
        double2 *xj = (double2 *)xb + (size_t)j * bt.k2c * p.mesh[0] * m2;
        if (bt.k2b == 0)
            e = wn_launch_contract(boxc, g->y3.p, g->e3.p, 1, bt.k3 + j, g->span[2], (int)m2, g->span[0] * g->span[1],
                                   g->span[2] * m2, m2, m2, 0, s0);
        if (e == hipSuccess)
            e = wn_launch_contract(g->y3.p, g->y23.p, g->e2.p, bt.k2c, bt.k2b, g->span[1], (int)m2, g->span[0],
                                   g->span[1] * m2, m2, (long long)bt.k2c * m2, m2, s0);
        if (e == hipSuccess)
            e = wn_launch_contract(g->y23.p, xj, g->e1.p, p.mesh[0], 0, g->span[0], (int)m2, bt.k2c, m2,
                                   (long long)bt.k2c * m2, (long long)p.mesh[0] * m2, m2, s0,
                                   g->ch32 ? g->sh.n * g->sh.n : 0);
    }
    return e;
}

// The kernel sequence of one planned mesh integral, ordered on g->stream (forked to g->stream2 and
// joined back).  Capture-safe: no allocation, no host synchronization, no memset.  marks (optional)
// holds WN_MARKS pre-created events per batch: 0 start, 1 k points, 2 interpolation, 8 eigensolver
// pass 0, 9 warm-start transform, 3 eigensolver done (stream 0); 7 start, 10 screening (SC / IC), 4 rotation,
// 5 response, 6 k sums (stream 1).
int enqueue_mesh(wnlqg_engine *g, std::vector<std::array<hipEvent_t, WN_MARKS>> *marks, double *interp_bytes,
                 bool overlap) {
    const MeshPlan &p = g->plan;
    const int nelem = output_elems(g);
    const long long nbatch = (long long)p.batches.size();
    WN_HIP(wn_launch_zero((double *)g->lanebuf.p, (long long)p.n_lanes * nelem * 2, g->stream));
    if (g->nterms) WN_HIP(wn_launch_zero((double *)g->lanecomp.p, (long long)p.n_lanes * nelem * 2, g->stream));
    WN_HIP(clear_flags(g));
    double2 *comp = g->nterms ? g->lanecomp.p : nullptr;
    const long long m2 = g->sh.ldo / 2;
    // overlap: rotation and response of batch b - 1 on the second stream while batch b is interpolated and
    // diagonalized (the captured batching graph and run_mesh's default: 0.91 of the one-stream time on the
    // 128^3 mesh); otherwise one stream, in program order, which times each kernel alone
    hipStream_t s0 = g->stream, s1 = overlap ? g->stream2 : g->stream;
    // fork the second stream from the first (this also brings it into a stream capture)
    WN_HIP(hipEventRecord(g->ev_fork, s0));
    WN_HIP(hipStreamWaitEvent(s1, g->ev_fork, 0));
    auto mark = [&](long long bi, int i, hipStream_t st) {
        if (marks) hipEventRecord((*marks)[bi][i], st);
    };
    for (long long bi = 0; bi < nbatch; ++bi) {
        const MeshBatch &bt = p.batches[bi];
        const long long k0 = bt.k0;
        const int cnt = bt.cnt;
        const int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        const int buf = (int)(bi & 1);
        double *xb = g->x.p + (size_t)buf * g->batch * g->sh.ldo;
        double *eb = g->e.p + (size_t)buf * g->batch * g->sh.n;
        double2 *ub = g->u.p + (size_t)buf * g->batch * g->sh.n * g->sh.n;
        // stream 0: k points, interpolation, eigensystem of batch bi into buffer bi % 2
        if (bi >= 2) WN_HIP(hipStreamWaitEvent(s0, g->ev_free[buf], 0));
        mark(bi, 0, s0);
        if (!p.sep) {
            WN_HIP(wn_launch_mesh_kpoints(g->kp.p, p.mesh, k0, cnt, rows, s0));
            WN_HIP(launch_phases(g, rows));
        }
        mark(bi, 1, s0);
        if (p.sep) {
            WN_HIP(launch_contract_batch(g, bt, xb, s0));
            if (interp_bytes) {
                // algorithmic bytes: the interpolated channels written (those after H in FP32 when ch32) and
                // the coefficient box read once per plane; the partial sums between stages are not counted
                const double nn = (double)g->sh.n * g->sh.n;
                *interp_bytes += (double)cnt * (g->ch32 ? 16.0 * nn + 8.0 * (m2 - nn) : 16.0 * m2);
                if (bt.k2b == 0)
                    *interp_bytes += (double)bt.nk3 * g->span[0] * g->span[1] * g->span[2] * m2 * 16.0;
            }
        } else {
            WN_HIP(launch_interp(g, rows, xb));
        }
        mark(bi, 2, s0);
        int *sweep_max = g->flags.p + WN_CTR_BASE;
        unsigned long long *sweep_total = (unsigned long long *)(g->flags.p + WN_CTR_BASE + 2);
        const int scheme = eig_scheme(g, bi);
        if (scheme == 2) {
            const double2 *uprev = g->u.p + (size_t)((bi - 1) & 1) * g->batch * g->sh.n * g->sh.n;
            mark(bi, 8, s0);
            WN_HIP(wn_launch_warm_transform_f64(xb, ub, cnt, g->sh, s0, uprev, g->escratch.p));
            mark(bi, 9, s0);
            WN_HIP(wn_launch_eig_pass_f64(xb, eb, ub, sweep_max, sweep_total, cnt, g->sh, g->eig_tpb, 2, s0,
                                          g->escratch.p, uprev, WN_CTR_STRIDE, g->flags.p));
        } else if (scheme == 1) {
            WN_HIP(wn_launch_eig_pass_f64(xb, eb, ub, sweep_max, sweep_total, cnt, g->sh, g->eig_tpb, 0, s0,
                                          g->escratch.p, nullptr, WN_CTR_STRIDE, g->flags.p));
            mark(bi, 8, s0);
            WN_HIP(wn_launch_warm_transform_f64(xb, ub, cnt, g->sh, s0, nullptr, g->escratch.p));
            mark(bi, 9, s0);
            WN_HIP(wn_launch_eig_pass_f64(xb, eb, ub, sweep_max, sweep_total, cnt, g->sh, g->eig_tpb, 1, s0,
                                          g->escratch.p, nullptr, WN_CTR_STRIDE, g->flags.p));
        } else {
            WN_HIP(wn_launch_eig_pass_f64(xb, eb, ub, sweep_max, sweep_total, cnt, g->sh, g->eig_tpb, -1, s0,
                                          g->escratch.p, nullptr, WN_CTR_STRIDE, g->flags.p));
        }
        mark(bi, 3, s0);
        WN_HIP(hipEventRecord(g->ev_ready[buf], s0));
        // stream 1: rotation, response and k sums of batch bi
        WN_HIP(hipStreamWaitEvent(s1, g->ev_ready[buf], 0));
        if (marks) hipEventRecord((*marks)[bi][7], s1);
        if (g->family == 0) {
            // SC / IC: screening first, so that the rotation skips points without pairs and reduces dd H and
            // dA to the pairs' elements
            WN_HIP(wn_launch_screen(eb, g->omega.p, g->count.p, g->pmeta.p, g->pmd.p, g->flags.p, cnt, g->sh, g->sc,
                                    s1));
            mark(bi, 10, s1);
            WN_HIP(wn_launch_rotate_sc(xb, ub, g->count.p, g->pmeta.p, g->pwd.p, cnt, g->sh, s1, g->rscratch.p,
                                       g->flags.p, g->ch32 && p.sep));
        } else {
            WN_HIP(wn_launch_rotate_f64(xb, ub, cnt, g->sh, s1, g->rscratch.p));
        }
        mark(bi, 4, s1);
        const int nch = (int)(p.bchunk[bi + 1] - p.bchunk[bi]);
        const int nln = (int)(p.blane[bi + 1] - p.blane[bi]);
        const int4 *chunks = p.dchunks.p + p.bchunk[bi];
        if (g->family == 0) {
            WN_HIP(wn_launch_pairs_f64(xb, eb, g->count.p, g->pmeta.p,
                                       wn_rotate_sc_elements(g->sh) ? g->pwd.p : nullptr, g->q.p, cnt, g->sh,
                                       g->sc, s1));
            mark(bi, 5, s1);
            WN_HIP(wn_launch_accum(chunks, nch, g->count.p, g->pmeta.p, g->pmd.p, g->q.p, g->omega.p, g->partial.p,
                                   g->sh, g->sc, s1));
        } else if (g->family == 3) {
            WN_HIP(wn_launch_shg(xb, eb, g->omega.p, chunks, 0, cnt, g->der.p, g->count.p, g->pidx.p, g->pdiff.p,
                                 g->coef.p, g->one.p, g->partial.p, g->sh, g->sp, s1));
            mark(bi, 5, s1);
            WN_HIP(wn_launch_shg(xb, eb, g->omega.p, chunks, nch, 0, g->der.p, g->count.p, g->pidx.p, g->pdiff.p,
                                 g->coef.p, g->one.p, g->partial.p, g->sh, g->sp, s1));
        } else {
            WN_HIP(wn_launch_geom_f64(xb, eb, g->omega.p, chunks, nch, nullptr, g->partial.p, g->flags.p, k0, g->sh, g->gp,
                                      s1));
            mark(bi, 5, s1);
        }
        WN_HIP(wn_launch_lane_reduce(p.dlanes.p + p.blane[bi], nln, g->partial.p, g->lanebuf.p, comp, nelem, s1));
        mark(bi, 6, s1);
        WN_HIP(hipEventRecord(g->ev_free[buf], s1));
    }
    // join
    WN_HIP(hipEventRecord(g->ev_join, s1));
    WN_HIP(hipStreamWaitEvent(s0, g->ev_join, 0));
    return WNLQG_OK;
}

// Lane sums to the caller after the enqueued integral: flags, prefactors, totals.  nk: the divisor of a
// mesh integral (N_k); 1 for a weighted point list, whose weights carry the normalization.
int finish_lanes(wnlqg_engine *g, int n_lanes, long long nk, double *lane_out, double *total_out) {
    const int nelem = output_elems(g);
    std::vector<double> lanes((size_t)n_lanes * nelem * 2);
    WN_HIP(hipMemcpyAsync(lanes.data(), g->lanebuf.p, lanes.size() * sizeof(double), hipMemcpyDeviceToHost,
                          g->stream));
    WN_HIP(hipStreamSynchronize(g->stream));
    int rc = check_flags(g);
    if (rc) return rc;
    if (g->family == 3) {
        // SHG: lane = raw sums (SecondHarmonicDriver.jl:178-185); total quantity q * factors[q] / N_k
        const size_t rel = (size_t)result_elems(g);
        for (int l = 0; l < n_lanes; ++l)
            shg_finalize(g, (const double2 *)(lanes.data() + (size_t)l * nelem * 2), lane_out + (size_t)l * rel * 2);
        for (size_t i = 0; i < (size_t)n_lanes * rel * 2; ++i)
            if (!std::isfinite(lane_out[i])) return fail(WNLQG_E_NONFINITE, "Nonfinite SHG response");
        if (total_out) {
            const double vol = std::fabs(g->cell_volume);
            const double factors[2] = {kCharge / (8.8541878128e-12 * vol) * 1e12, kCharge * kCharge / (kHbarJs * vol)};
            for (size_t i = 0; i < rel * 2; ++i) total_out[i] = 0.0;
            for (int l = 0; l < n_lanes; ++l)
                for (size_t i = 0; i < rel * 2; ++i) total_out[i] += lane_out[(size_t)l * rel * 2 + i];
            // column-major [nw, nc, 7, 2]: quantity q occupies complex entries [q rel/2, (q+1) rel/2)
            for (size_t i = 0; i < rel; ++i) {
                const double f = factors[i < rel / 2 ? 0 : 1] / (double)nk;
                total_out[2 * i] *= f;
                total_out[2 * i + 1] *= f;
            }
        }
        return WNLQG_OK;
    }
    // SC / IC: each state * its prefactor, then 1/N_k (consume_point!, IntegralDriver.jl:820-874).
    // BCD / AHC: lanes hold raw sums of the per-point values (BerryMultipoleDriver.jl:279-285); the total is
    // divided by N_k and by the measure (BerryMultipoleDriver.jl:290-303, SpectralResponseDriver.jl:541-552).
    const double scale = g->family ? 1.0 : 1.0 / (double)nk;
    for (int l = 0; l < n_lanes; ++l)
        for (int e = 0; e < nelem; ++e) {
            const size_t i = ((size_t)l * nelem + e) * 2;
            lane_out[i] = (g->elem_pref[e] * lanes[i]) * scale;
            lane_out[i + 1] = (g->elem_pref[e] * lanes[i + 1]) * scale;
        }
    for (size_t i = 0; i < lanes.size(); ++i)
        if (!std::isfinite(lane_out[i])) return fail(WNLQG_E_NONFINITE, "non-finite response value");
    if (total_out) {
        for (size_t i = 0; i < (size_t)nelem * 2; ++i) total_out[i] = 0.0;
        for (int l = 0; l < n_lanes; ++l)
            for (size_t i = 0; i < (size_t)nelem * 2; ++i) total_out[i] += lane_out[(size_t)l * nelem * 2 + i];
        for (int e = 0; e < nelem; ++e)
            if (g->elem_norm[e] != 0.0) {
                total_out[2 * e] = total_out[2 * e] / (double)nk / g->elem_norm[e];
                total_out[2 * e + 1] = total_out[2 * e + 1] / (double)nk / g->elem_norm[e];
            }
    }
    return WNLQG_OK;
}

}  // namespace

extern "C" {

int32_t wnlqg_run_mesh(wnlqg_engine *g, const int32_t mesh[3], int32_t n_lanes, double *lane_out, double *total_out,
                       wnlqg_stats *stats) {
    if (!g || !mesh || !lane_out) return fail(WNLQG_E_INVALID, "invalid arguments");
    if (mesh[0] < 1 || mesh[1] < 1 || mesh[2] < 1) return fail(WNLQG_E_INVALID, "mesh sizes must be positive");
    const long long nk = (long long)mesh[0] * mesh[1] * mesh[2];
    if (n_lanes < 1 || n_lanes > nk) return fail(WNLQG_E_INVALID, "n_lanes must be in [1, N_k]");
    WN_HIP(hipSetDevice(g->device));
    int rc = plan_mesh(g, mesh, n_lanes);
    if (rc) return rc;
    const long long nbatch = (long long)g->plan.batches.size();
    std::vector<std::array<hipEvent_t, WN_MARKS>> marks(nbatch);
    bool events_ok = true;
    for (auto &mk : marks)
        for (auto &e : mk) {
            e = nullptr;
            if (events_ok && hipEventCreate(&e) != hipSuccess) events_ok = false;
        }
    auto release_events = [&]() {
        for (auto &mk : marks)
            for (auto e : mk)
                if (e) hipEventDestroy(e);
    };
    double interp_bytes = 0.0;
    // the two-stream order of the batching graph (rotation and response of batch b - 1 beside the
    // interpolation and eigensystem of batch b: the stages are bound by different units, HBM writes, FP64
    // VALU and barrier latency), whose stage timers span shared GPU time; WNLQG_MESH_OVERLAP=0 selects one
    // stream in program order, each kernel then timed alone
    bool overlap = true;
    if (const char *ov = std::getenv("WNLQG_MESH_OVERLAP")) overlap = strcmp(ov, "0") != 0;
    rc = enqueue_mesh(g, events_ok ? &marks : nullptr, &interp_bytes, overlap);
    if (!rc) rc = finish_lanes(g, g->plan.n_lanes, g->plan.nk, lane_out, total_out);
    if (rc) {
        release_events();
        return rc;
    }
    wnlqg_stats st{};
    float ms = 0.0f;
    if (events_ok) {
        for (long long bi = 0; bi < nbatch; ++bi) {
            auto &mk = marks[bi];
            hipEventElapsedTime(&ms, mk[1], mk[2]); st.interp_ms += ms;
            const int scheme = eig_scheme(g, bi);
            if (scheme) {
                hipEventElapsedTime(&ms, mk[2], mk[8]); st.eig_ms += ms;
                hipEventElapsedTime(&ms, mk[8], mk[9]); st.warm_ms += ms;
                hipEventElapsedTime(&ms, mk[9], mk[3]); st.eig_ms += ms;
                st.eig_launches += scheme == 1 ? 2 : 1;
            } else {
                hipEventElapsedTime(&ms, mk[2], mk[3]); st.eig_ms += ms;
                st.eig_launches += 1;
            }
            if (g->family == 0) {
                // screening counts with the response kernels; the rotation timer holds k_rotate_sc alone
                hipEventElapsedTime(&ms, mk[7], mk[10]); st.pairs_ms += ms;
                hipEventElapsedTime(&ms, mk[10], mk[4]); st.rotate_ms += ms;
            } else {
                hipEventElapsedTime(&ms, mk[7], mk[4]); st.rotate_ms += ms;
            }
            hipEventElapsedTime(&ms, mk[4], mk[5]); st.pairs_ms += ms;
            hipEventElapsedTime(&ms, mk[5], mk[6]); st.accum_ms += ms;
        }
        if (nbatch > 0) {
            hipEventElapsedTime(&ms, marks[0][0], marks[nbatch - 1][6]);
            st.total_ms = ms;
        }
    }
    release_events();
    std::vector<int> fl(WN_FLAG_WORDS);
    WN_HIP(hipMemcpy(fl.data(), g->flags.p, fl.size() * sizeof(int), hipMemcpyDeviceToHost));
    unsigned long long swt = 0, npairs = 0;
    st.max_sweeps = 0;
    for (int b = 0; b < WN_CTR_BUCKETS; ++b) {
        const int *bk = fl.data() + WN_CTR_BASE + WN_CTR_STRIDE * b;
        unsigned long long v = 0;
        st.max_sweeps = std::max(st.max_sweeps, bk[0]);
        std::memcpy(&v, bk + 2, sizeof(v));
        swt += v;
        std::memcpy(&v, bk + 4, sizeof(v));
        npairs += v;
        st.rotate_element_points += bk[6];
        st.rotate_full_points += bk[7];
        st.rotate_element_bands += bk[8];
        st.rotate_row_points += bk[9];
    }
    st.mean_sweeps = nk > 0 ? (double)swt / (double)nk : 0.0;
    st.active_pairs = (int64_t)npairs;
    st.kpoints = nk;
    st.batches = (int)nbatch;
    st.interp_launches = (int)nbatch;
    st.interp_flops = 8.0 * g->sh.nr * (double)g->sh.n * g->sh.n * g->sh.C * (double)nk;
    st.interp_bytes = interp_bytes;
    st.separable = g->plan.sep ? 1 : 0;
    if (stats) *stats = st;
    return WNLQG_OK;
}

// Weighted point list (SC / IC): the k-loop of a symmetry-reduced mesh, whose evaluation items are the
// orbit representatives with weight multiplicity / N_k (consume_point!, IntegralDriver.jl:822-831).  Per
// batch: interpolation and eigensystem of the listed points, screening, weights folded into the pairs'
// occupation differences, rotation and pair kernels, frequency contraction per lane-aligned chunk, lane
// sums in chunk order.
int32_t wnlqg_run_points(wnlqg_engine *g, const double *kpoints, const double *weights, int64_t m, int32_t n_lanes,
                         double *lane_out, double *total_out) {
    if (!g || !kpoints || !weights || !lane_out || m < 1) return fail(WNLQG_E_INVALID, "invalid arguments");
    if (n_lanes < 1 || n_lanes > m) return fail(WNLQG_E_INVALID, "n_lanes must be in [1, m]");
    if (g->family != 0 && g->family != 1)
        return fail(WNLQG_E_INVALID, "weighted point lists integrate shift / injection current, BCD and AHC engines only");
    if (g->backend == WNLQG_FOURIER_MIXED)
        return fail(WNLQG_E_UNSUPPORTED, "weighted k-point lists (adaptive / region nodes) support direct evaluation; "
                    "use fourier_backend direct or auto");
    if (g->kpc < 1 || g->kpc > WN_ACC_KMAX) return fail(WNLQG_E_INVALID, "chunk size %d outside [1, %d]", g->kpc, WN_ACC_KMAX);
    for (int64_t i = 0; i < m; ++i)
        if (!std::isfinite(weights[i])) return fail(WNLQG_E_NONFINITE, "non-finite weight %lld", (long long)i);
    WN_HIP(hipSetDevice(g->device));
    const int nelem = output_elems(g);
    if (g->lanebuf.n < (size_t)n_lanes * nelem) {
        WN_HIP(g->lanebuf.alloc((size_t)n_lanes * nelem));
        WN_HIP(g->lanecomp.alloc((size_t)n_lanes * nelem));
    }
    if (g->wts.n < (size_t)g->batch) WN_HIP(g->wts.alloc((size_t)g->batch));
    WN_HIP(wn_launch_zero((double *)g->lanebuf.p, (long long)n_lanes * nelem * 2, g->stream));
    if (g->nterms) WN_HIP(wn_launch_zero((double *)g->lanecomp.p, (long long)n_lanes * nelem * 2, g->stream));
    WN_HIP(clear_flags(g));
    double2 *comp = g->nterms ? g->lanecomp.p : nullptr;
    std::vector<long long> lstart(n_lanes + 1);
    for (int l = 0; l <= n_lanes; ++l) lstart[l] = (long long)((__int128)l * m / n_lanes);
    std::vector<int4> ch, ln;
    for (long long p0 = 0; p0 < m; p0 += g->batch) {
        const int cnt = (int)std::min<long long>(g->batch, m - p0);
        const int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        // chunks of at most kpc points inside each lane part of this batch, in point order
        ch.clear();
        ln.clear();
        int l = (int)(std::upper_bound(lstart.begin(), lstart.end(), p0) - lstart.begin()) - 1;
        for (; l < n_lanes && lstart[l] < p0 + cnt; ++l) {
            const long long a = std::max(p0, lstart[l]), b = std::min(p0 + cnt, lstart[l + 1]);
            if (a >= b) continue;
            const int cb = (int)ch.size();
            for (long long c = a; c < b; c += g->kpc)
                ch.push_back(make_int4((int)(c - p0), (int)std::min<long long>(g->kpc, b - c), l, 0));
            ln.push_back(make_int4(l, cb, (int)ch.size(), 0));
        }
        // upload_points drains the stream first, so the previous batch no longer reads kp, wts, chunks, lanes
        int rc = upload_points(g, kpoints + 3 * (size_t)p0, cnt, rows);
        if (rc) return rc;
        rc = ensure_partial(g, (int)ch.size());
        if (rc) return rc;
        if (g->lanes.n < ln.size()) WN_HIP(g->lanes.alloc(std::max<size_t>(ln.size(), (size_t)n_lanes)));
        WN_HIP(hipMemcpy(g->wts.p, weights + p0, (size_t)cnt * sizeof(double), hipMemcpyHostToDevice));
        WN_HIP(hipMemcpy(g->chunks.p, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice));
        WN_HIP(hipMemcpy(g->lanes.p, ln.data(), ln.size() * sizeof(int4), hipMemcpyHostToDevice));
        rc = run_front(g, cnt, rows, true);
        if (rc) return rc;
        if (g->family == 1) {
            // BCD / AHC: every channel rotated, each point's moments scaled by its weight inside the chunk sums
            WN_HIP(wn_launch_rotate_f64(g->x.p, g->u.p, cnt, g->sh, g->stream, g->rscratch.p));
            WN_HIP(wn_launch_geom_f64(g->x.p, g->e.p, g->omega.p, g->chunks.p, (int)ch.size(), g->wts.p, g->partial.p,
                                      g->flags.p, p0, g->sh, g->gp, g->stream));
            WN_HIP(wn_launch_lane_reduce(g->lanes.p, (int)ln.size(), g->partial.p, g->lanebuf.p, comp, nelem,
                                         g->stream));
            continue;
        }
        WN_HIP(wn_launch_screen(g->e.p, g->omega.p, g->count.p, g->pmeta.p, g->pmd.p, g->flags.p, cnt, g->sh, g->sc,
                                g->stream));
        WN_HIP(wn_launch_weight_pairs(g->pmd.p, g->count.p, g->wts.p, cnt, g->sh, g->stream));
        WN_HIP(wn_launch_rotate_sc(g->x.p, g->u.p, g->count.p, g->pmeta.p, g->pwd.p, cnt, g->sh, g->stream,
                                   g->rscratch.p, g->flags.p));
        WN_HIP(wn_launch_pairs_f64(g->x.p, g->e.p, g->count.p, g->pmeta.p,
                                   wn_rotate_sc_elements(g->sh) ? g->pwd.p : nullptr, g->q.p, cnt, g->sh, g->sc,
                                   g->stream));
        WN_HIP(wn_launch_accum(g->chunks.p, (int)ch.size(), g->count.p, g->pmeta.p, g->pmd.p, g->q.p, g->omega.p,
                               g->partial.p, g->sh, g->sc, g->stream));
        WN_HIP(wn_launch_lane_reduce(g->lanes.p, (int)ln.size(), g->partial.p, g->lanebuf.p, comp, nelem, g->stream));
    }
    return finish_lanes(g, n_lanes, 1, lane_out, total_out);
}

// Symmetry-reduced mesh (SC / IC, separable interpolation): the first two contraction stages of every
// k3 plane run whole (they are shared by all points of the plane), the last stage writes only the orbit
// representatives' rows, compactly, and only they are diagonalized (warm-started in groups of 4 consecutive
// representatives where the size allows, as the mesh rows are), screened,
// weighted, rotated and contracted.  reps: sorted linear mesh indices (axis 1 fastest); weights: the
// items' scales (multiplicity / N_k); lanes partition the m items as in wnlqg_run_points.
int32_t wnlqg_run_mesh_reduced(wnlqg_engine *g, const int32_t mesh[3], const int64_t *reps, const double *weights,
                               int64_t m, int32_t n_lanes, double *lane_out, double *total_out) {
    if (!g || !mesh || !reps || !weights || !lane_out || m < 1) return fail(WNLQG_E_INVALID, "invalid arguments");
    if (mesh[0] < 1 || mesh[1] < 1 || mesh[2] < 1) return fail(WNLQG_E_INVALID, "mesh sizes must be positive");
    if (n_lanes < 1 || n_lanes > m) return fail(WNLQG_E_INVALID, "n_lanes must be in [1, m]");
    if (g->family != 0)
        return fail(WNLQG_E_INVALID, "reduced meshes integrate shift / injection current engines only");
    const long long nk = (long long)mesh[0] * mesh[1] * mesh[2];
    if (m > nk) return fail(WNLQG_E_INVALID, "more representatives than mesh points");
    for (int64_t i = 0; i < m; ++i) {
        if (reps[i] < 0 || reps[i] >= nk || (i > 0 && reps[i] <= reps[i - 1]))
            return fail(WNLQG_E_INVALID, "representatives must be strictly increasing mesh indices");
        if (!std::isfinite(weights[i])) return fail(WNLQG_E_NONFINITE, "non-finite weight %lld", (long long)i);
    }
    WN_HIP(hipSetDevice(g->device));
    int rc = plan_mesh(g, mesh, 1);
    if (rc) return rc;
    const MeshPlan &p = g->plan;
    if (!p.sep)
        return fail(WNLQG_E_INVALID, "reduced meshes need the separable interpolation (fourier_backend auto, planes or mixed)");
    const long long nbatch = (long long)p.batches.size();
    // representatives of each batch: a contiguous range of the sorted list
    std::vector<long long> rb(nbatch + 1, m);
    for (long long bi = 0; bi < nbatch; ++bi)
        rb[bi] = std::lower_bound(reps, reps + m, (int64_t)p.batches[bi].k0) - reps;
    // per representative its k1 (the phase row of the last contraction stage); per (batch, plane) the runs
    // of representatives sharing a mesh row: (row b within the plane's k2 range, first, end) with indices
    // relative to the batch's first representative
    std::vector<int> sel(m);
    std::vector<int4> groups;
    std::vector<long long> gplane;   // groups of plane j of batch bi: [gplane[gp0[bi] + j], gplane[gp0[bi] + j + 1])
    std::vector<long long> gp0(nbatch + 1, 0);
    for (long long bi = 0; bi < nbatch; ++bi) {
        const MeshBatch &bt = p.batches[bi];
        const long long rowsz = (long long)bt.k2c * mesh[0];
        gp0[bi] = (long long)gplane.size();
        long long i = rb[bi];
        for (int j = 0; j < bt.nk3; ++j) {
            gplane.push_back((long long)groups.size());
            while (i < rb[bi + 1] && (reps[i] - bt.k0) / rowsz == j) {
                const long long r = (reps[i] - bt.k0) % rowsz;
                const int b = (int)(r / mesh[0]);
                const long long t0 = i;
                while (i < rb[bi + 1] && (reps[i] - bt.k0) / rowsz == j && ((reps[i] - bt.k0) % rowsz) / mesh[0] == b) {
                    sel[i] = (int)((reps[i] - bt.k0) % mesh[0]);
                    ++i;
                }
                groups.push_back(make_int4(b, (int)(t0 - rb[bi]), (int)(i - rb[bi]), 0));
            }
        }
        gplane.push_back((long long)groups.size());
    }
    gp0[nbatch] = (long long)gplane.size();
    std::vector<long long> lstart(n_lanes + 1);
    for (int l = 0; l <= n_lanes; ++l) lstart[l] = (long long)((__int128)l * m / n_lanes);
    std::vector<int4> ch, ln;
    std::vector<long long> bch(nbatch + 1, 0), bln(nbatch + 1, 0);
    int max_chunks = 0;
    for (long long bi = 0; bi < nbatch; ++bi) {
        bch[bi] = (long long)ch.size();
        bln[bi] = (long long)ln.size();
        const long long r0 = rb[bi], r1 = rb[bi + 1];
        int l = (int)(std::upper_bound(lstart.begin(), lstart.end(), r0) - lstart.begin()) - 1;
        for (; r0 < r1 && l < n_lanes && lstart[l] < r1; ++l) {
            const long long a = std::max(r0, lstart[l]), b = std::min(r1, lstart[l + 1]);
            if (a >= b) continue;
            const int cb = (int)(ch.size() - bch[bi]);
            for (long long c = a; c < b; c += g->kpc)
                ch.push_back(make_int4((int)(c - r0), (int)std::min<long long>(g->kpc, b - c), l, 0));
            ln.push_back(make_int4(l, cb, (int)(ch.size() - bch[bi]), 0));
        }
        max_chunks = std::max(max_chunks, (int)(ch.size() - bch[bi]));
    }
    bch[nbatch] = (long long)ch.size();
    bln[nbatch] = (long long)ln.size();
    rc = ensure_partial(g, std::max(1, max_chunks));
    if (rc) return rc;
    const int nelem = output_elems(g);
    if (g->lanebuf.n < (size_t)n_lanes * nelem) {
        WN_HIP(g->lanebuf.alloc((size_t)n_lanes * nelem));
        WN_HIP(g->lanecomp.alloc((size_t)n_lanes * nelem));
    }
    if (g->wts.n < (size_t)m) WN_HIP(g->wts.alloc((size_t)m));
    if (g->ridxs.n < (size_t)m) WN_HIP(g->ridxs.alloc((size_t)m));
    if (g->rgroups.n < groups.size()) WN_HIP(g->rgroups.alloc(groups.size()));
    if (g->rchunks.n < ch.size()) WN_HIP(g->rchunks.alloc(ch.size()));
    if (g->rlanes.n < ln.size()) WN_HIP(g->rlanes.alloc(ln.size()));
    WN_HIP(hipStreamSynchronize(g->stream));
    WN_HIP(hipMemcpy(g->wts.p, weights, (size_t)m * sizeof(double), hipMemcpyHostToDevice));
    WN_HIP(hipMemcpy(g->ridxs.p, sel.data(), (size_t)m * sizeof(int), hipMemcpyHostToDevice));
    WN_HIP(hipMemcpy(g->rgroups.p, groups.data(), groups.size() * sizeof(int4), hipMemcpyHostToDevice));
    WN_HIP(hipMemcpy(g->rchunks.p, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice));
    WN_HIP(hipMemcpy(g->rlanes.p, ln.data(), ln.size() * sizeof(int4), hipMemcpyHostToDevice));
    hipStream_t s0 = g->stream;
    WN_HIP(wn_launch_zero((double *)g->lanebuf.p, (long long)n_lanes * nelem * 2, s0));
    if (g->nterms) WN_HIP(wn_launch_zero((double *)g->lanecomp.p, (long long)n_lanes * nelem * 2, s0));
    WN_HIP(clear_flags(g));
    double2 *comp = g->nterms ? g->lanecomp.p : nullptr;
    int *sweep_max = g->flags.p + WN_CTR_BASE;
    unsigned long long *sweep_total = (unsigned long long *)(g->flags.p + WN_CTR_BASE + 2);
    // the representatives' rows, compact, alternately in the two halves of the double buffer, so that a
    // batch whose representatives sit at the same places of the previous batch's plane (point groups that
    // keep k3) starts its eigensolver from the previous plane's eigenvectors, as the full mesh does
    const long long m2 = g->sh.ldo / 2;
    const double2 *boxc = (const double2 *)g->box.p;
    long long prev = -1;
    int buf = 0;
    auto same_pattern = [&](long long a, long long b) {
        const MeshBatch &ba = p.batches[a], &bb = p.batches[b];
        if (ba.nk3 != 1 || bb.nk3 != 1 || ba.k2c != mesh[1] || bb.k2c != mesh[1]) return false;
        if (rb[a + 1] - rb[a] != rb[b + 1] - rb[b]) return false;
        for (long long i = 0; i < rb[a + 1] - rb[a]; ++i)
            if (reps[rb[a] + i] - ba.k0 != reps[rb[b] + i] - bb.k0) return false;
        return true;
    };
    for (long long bi = 0; bi < nbatch; ++bi) {
        const int cr = (int)(rb[bi + 1] - rb[bi]);
        const MeshBatch &bt = p.batches[bi];
        double *xo = g->x.p + (size_t)buf * g->batch * g->sh.ldo;
        double *eo = g->e.p + (size_t)buf * g->batch * g->sh.n;
        double2 *uo = g->u.p + (size_t)buf * g->batch * g->sh.n * g->sh.n;
        for (int j = 0; j < bt.nk3; ++j) {
            // the first stage whenever a plane starts (batches holding the rest of the plane reuse it), the
            // second for the batch's k2 range, the last only for the representatives' rows
            if (bt.k2b == 0)
                WN_HIP(wn_launch_contract(boxc, g->y3.p, g->e3.p, 1, bt.k3 + j, g->span[2], (int)m2,
                                          g->span[0] * g->span[1], g->span[2] * m2, m2, m2, 0, s0));
            const long long ga = gplane[gp0[bi] + j], gb = gplane[gp0[bi] + j + 1];
            if (ga == gb) continue;
            WN_HIP(wn_launch_contract(g->y3.p, g->y23.p, g->e2.p, bt.k2c, bt.k2b, g->span[1], (int)m2, g->span[0],
                                      g->span[1] * m2, m2, (long long)bt.k2c * m2, m2, s0));
            WN_HIP(wn_launch_contract_sel(g->y23.p, (double2 *)xo, g->e1.p, g->rgroups.p + ga, (int)(gb - ga),
                                          g->ridxs.p + rb[bi], g->span[0], (int)m2, m2, (long long)bt.k2c * m2, m2,
                                          s0));
        }
        if (cr == 0) continue;
        if (wn_warm_start(g->sh.n, cr) && prev >= 0 && prev == bi - 1 && same_pattern(prev, bi)) {
            const double2 *uprev = g->u.p + (size_t)(buf ^ 1) * g->batch * g->sh.n * g->sh.n;
            WN_HIP(wn_launch_warm_transform_f64(xo, uo, cr, g->sh, s0, uprev, g->escratch.p));
            WN_HIP(wn_launch_eig_pass_f64(xo, eo, uo, sweep_max, sweep_total, cr, g->sh, g->eig_tpb, 2, s0, g->escratch.p,
                                          uprev, WN_CTR_STRIDE, g->flags.p));
        } else if (wn_warm_start(g->sh.n, cr)) {
            // groups of 4 consecutive representatives (mostly neighbours along axis 1)
            WN_HIP(wn_launch_eig_pass_f64(xo, eo, uo, sweep_max, sweep_total, cr, g->sh, g->eig_tpb, 0, s0, g->escratch.p,
                                          nullptr, WN_CTR_STRIDE, g->flags.p));
            WN_HIP(wn_launch_warm_transform_f64(xo, uo, cr, g->sh, s0, nullptr, g->escratch.p));
            WN_HIP(wn_launch_eig_pass_f64(xo, eo, uo, sweep_max, sweep_total, cr, g->sh, g->eig_tpb, 1, s0, g->escratch.p,
                                          nullptr, WN_CTR_STRIDE, g->flags.p));
        } else {
            WN_HIP(wn_launch_eig_pass_f64(xo, eo, uo, sweep_max, sweep_total, cr, g->sh, g->eig_tpb, -1, s0,
                                          g->escratch.p, nullptr, WN_CTR_STRIDE, g->flags.p));
        }
        WN_HIP(wn_launch_screen(eo, g->omega.p, g->count.p, g->pmeta.p, g->pmd.p, g->flags.p, cr, g->sh, g->sc, s0));
        WN_HIP(wn_launch_weight_pairs(g->pmd.p, g->count.p, g->wts.p + rb[bi], cr, g->sh, s0));
        WN_HIP(wn_launch_rotate_sc(xo, uo, g->count.p, g->pmeta.p, g->pwd.p, cr, g->sh, s0, g->rscratch.p,
                                   g->flags.p));
        WN_HIP(wn_launch_pairs_f64(xo, eo, g->count.p, g->pmeta.p, wn_rotate_sc_elements(g->sh) ? g->pwd.p : nullptr,
                                   g->q.p, cr, g->sh, g->sc, s0));
        WN_HIP(wn_launch_accum(g->rchunks.p + bch[bi], (int)(bch[bi + 1] - bch[bi]), g->count.p, g->pmeta.p, g->pmd.p,
                               g->q.p, g->omega.p, g->partial.p, g->sh, g->sc, s0));
        WN_HIP(wn_launch_lane_reduce(g->rlanes.p + bln[bi], (int)(bln[bi + 1] - bln[bi]), g->partial.p, g->lanebuf.p,
                                     comp, nelem, s0));
        prev = bi;
        buf ^= 1;
    }
    return finish_lanes(g, n_lanes, 1, lane_out, total_out);
}

// ------------------------------------------------------------------------------------------------
// Batching form: many (model, task, mesh) triples of uneven n_w in one captured HIP graph.
// ------------------------------------------------------------------------------------------------
}  // extern "C"

struct wnlqg_batch {
    int device = 0;
    std::vector<wnlqg_engine *> engines;
    hipStream_t s0 = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// complex channels stored per k point for a task family (H, dH, ddH, A, dA as needed)
int channels_for(int family, bool pos) {
    if (family == 1) return pos ? 16 : 4;
    return pos ? WN_CH_FULL : WN_CH_BASE;
}

// Autotuner: eigensolver form for each distinct n_w, timed on the first batch of each size class's first
// engine (interpolation of that batch, then the register-resident-V kernel and the LDS kernel with 64, 128,
// 256 threads per k point).
// n_w > 64 has a single eigensolver form (512 threads, V in global memory) and is not tuned.  Only
// configurations whose launch and timing succeeded are eligible.
int autotune(wnlqg_batch *b) {
    std::vector<int> done(129, 0);
    hipEvent_t a, z;
    WN_HIP(hipEventCreate(&a));
    if (hipEventCreate(&z) != hipSuccess) {
        hipEventDestroy(a);
        return fail(WNLQG_E_HIP, "autotune event creation failed");
    }
    int rc = WNLQG_OK;
    for (wnlqg_engine *g : b->engines) {
        const int n = g->sh.n;
        if (n < 1 || n > 64) continue;
        if (done[n]) {
            g->eig_tpb = done[n];
            continue;
        }
        const int cnt = (int)std::min<long long>(std::min<long long>(g->batch, 8192), g->plan.nk);
        const int rows = (cnt + WN_GM - 1) / WN_GM * WN_GM;
        if (wn_launch_mesh_kpoints(g->kp.p, g->plan.mesh, 0, cnt, rows, g->stream) != hipSuccess) {
            rc = fail(WNLQG_E_HIP, "autotune probe failed");
            break;
        }
        rc = run_front(g, cnt, rows, false);
        if (rc) break;
        float best = 1e30f;
        int best_tpb = 0;
        // candidates: the deferred-eigenvector and register-resident-V kernels (n_w <= 32), then the LDS kernel at
        // 64, 128, 256 threads
        std::vector<int> cands;
        if (wn_eig_rv_supported(n)) {
            cands.push_back(WN_EIG_DV);
            cands.push_back(WN_EIG_RV);
        }
        for (int tpb : {64, 128, 256}) cands.push_back(tpb);
        for (int tpb : cands) {
            float ms = 0.0f;
            bool ok = true;
            for (int rep = 0; ok && rep < 2; ++rep) {  // first launch warms up
                ok = hipEventRecord(a, g->stream) == hipSuccess &&
                     wn_launch_eig_f64(g->x.p, g->e.p, g->u.p, g->sweeps.p, cnt, g->sh, tpb, g->stream, g->escratch.p,
                                       g->flags.p) == hipSuccess &&
                     hipEventRecord(z, g->stream) == hipSuccess;
                if (hipEventSynchronize(z) != hipSuccess) {
                    rc = fail(WNLQG_E_HIP, "autotune timing failed");
                    break;
                }
                ok = ok && hipEventElapsedTime(&ms, a, z) == hipSuccess;
            }
            if (rc) break;
            if (ok && ms < best) {
                best = ms;
                best_tpb = tpb;
            }
        }
        if (rc) break;
        g->eig_tpb = best_tpb;
        done[n] = best_tpb;
    }
    hipEventDestroy(a);
    hipEventDestroy(z);
    return rc;
}

}  // namespace

extern "C" {

int32_t wnlqg_batch_destroy(wnlqg_batch *b) {
    if (!b) return WNLQG_OK;
    (void)hipSetDevice(b->device);
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto e : b->events) (void)hipEventDestroy(e);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    for (auto st : b->streams) (void)hipStreamDestroy(st);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    for (auto g : b->engines) wnlqg_destroy(g);
    delete b;
    return WNLQG_OK;
}

int32_t wnlqg_batch_create(int32_t count, const wnlqg_model *models, const wnlqg_task *tasks, const int32_t *meshes,
                           const int32_t *n_lanes, int32_t autotune_tiles, wnlqg_batch **out) {
    if (!out) return fail(WNLQG_E_INVALID, "out is NULL");
    *out = nullptr;
    if (count < 1 || !models || !tasks || !meshes || !n_lanes) return fail(WNLQG_E_INVALID, "invalid arguments");
    for (int i = 0; i < count; ++i) {
        const int32_t *m = meshes + 3 * i;
        if (m[0] < 1 || m[1] < 1 || m[2] < 1) return fail(WNLQG_E_INVALID, "entry %d: mesh sizes must be positive", i);
        const long long nk = (long long)m[0] * m[1] * m[2];
        if (n_lanes[i] < 1 || n_lanes[i] > nk) return fail(WNLQG_E_INVALID, "entry %d: n_lanes must be in [1, N_k]", i);
    }
    wnlqg_batch *b = new wnlqg_batch();
    auto bail = [&](int code) {
        wnlqg_batch_destroy(b);
        return code;
    };
    if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(WNLQG_E_HIP, "no HIP device"));
    // one engine per entry; working sets share a 96 GB budget unless the task fixes its batch size
    for (int i = 0; i < count; ++i) {
        wnlqg_task t = tasks[i];
        if (t.batch_size == 0) {
            const int fam = task_family(t.quantities);
            const double n2 = (double)models[i].num_orbitals * models[i].num_orbitals;
            const double per_k = 2.4 * channels_for(fam, models[i].position_r != nullptr) * n2 * 8.0 + 64.0 * n2 + 4096.0;
            const long long nk = (long long)meshes[3 * i] * meshes[3 * i + 1] * meshes[3 * i + 2];
            long long bs = (long long)(96.0e9 / count / per_k);
            bs = std::max(128LL, std::min({bs, 65536LL, std::max(nk, 128LL)}));
            t.batch_size = (int32_t)bs;
        }
        wnlqg_engine *g = nullptr;
        int rc = wnlqg_create(models + i, &t, &g);
        if (rc) return bail(fail(rc, "entry %d: %s", i, g_last_error.c_str()));
        b->engines.push_back(g);
        rc = plan_mesh(g, meshes + 3 * i, n_lanes[i]);
        if (rc) return bail(rc);
        if (hipStreamSynchronize(g->stream) != hipSuccess) return bail(fail(WNLQG_E_HIP, "entry %d: setup failed", i));
    }
    if (autotune_tiles) {
        int rc = autotune(b);
        if (rc) return bail(rc);
    }
    // capture every entry's kernel sequence into one graph: entries are dealt round-robin to four
    // streams forked from and joined back into the capture stream, so independent models overlap
    const int ns = std::min(count, 4);
    if (hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess)
        return bail(fail(WNLQG_E_HIP, "stream / event creation failed"));
    for (int s = 0; s < ns; ++s) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(WNLQG_E_HIP, "stream creation failed"));
        b->streams.push_back(st);
    }
    for (int e = 0; e < ns + 1; ++e) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return bail(fail(WNLQG_E_HIP, "event creation failed"));
        b->events.push_back(ev);
    }
    if (hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bail(fail(WNLQG_E_HIP, "stream capture failed to start"));
    int rc = WNLQG_OK;
    hipError_t he = hipEventRecord(b->events[0], b->s0);
    for (int s = 0; s < ns && he == hipSuccess; ++s) he = hipStreamWaitEvent(b->streams[s], b->events[0], 0);
    for (int i = 0; i < count && he == hipSuccess && !rc; ++i) {
        wnlqg_engine *g = b->engines[i];
        hipStream_t own = g->stream;
        g->stream = b->streams[i % ns];
        rc = enqueue_mesh(g, nullptr, nullptr, true);
        g->stream = own;
    }
    for (int s = 0; s < ns && he == hipSuccess; ++s) {
        he = hipEventRecord(b->events[s + 1], b->streams[s]);
        if (he == hipSuccess) he = hipStreamWaitEvent(b->s0, b->events[s + 1], 0);
    }
    hipError_t ce = hipStreamEndCapture(b->s0, &b->graph);
    if (rc) return bail(rc);
    if (he != hipSuccess || ce != hipSuccess) return bail(fail(WNLQG_E_HIP, "graph capture failed"));
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        return bail(fail(WNLQG_E_HIP, "graph instantiation failed"));
    *out = b;
    return WNLQG_OK;
}

int32_t wnlqg_batch_run(wnlqg_batch *b, double *const *lane_out, double *const *total_out, double *elapsed_ms) {
    if (!b || !lane_out) return fail(WNLQG_E_INVALID, "invalid arguments");
    for (size_t i = 0; i < b->engines.size(); ++i)
        if (!lane_out[i]) return fail(WNLQG_E_INVALID, "lane_out[%zu] is NULL", i);
    WN_HIP(hipSetDevice(b->device));
    WN_HIP(hipEventRecord(b->t0, b->s0));
    WN_HIP(hipGraphLaunch(b->exec, b->s0));
    WN_HIP(hipEventRecord(b->t1, b->s0));
    WN_HIP(hipStreamSynchronize(b->s0));
    float ms = 0.0f;
    WN_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    if (elapsed_ms) *elapsed_ms = ms;
    for (size_t i = 0; i < b->engines.size(); ++i) {
        int rc = finish_lanes(b->engines[i], b->engines[i]->plan.n_lanes, b->engines[i]->plan.nk, lane_out[i],
                              total_out ? total_out[i] : nullptr);
        if (rc) return fail(rc, "entry %zu: %s", i, g_last_error.c_str());
    }
    return WNLQG_OK;
}

int32_t wnlqg_batch_tuning(const wnlqg_batch *b, int32_t index, int32_t *eig_threads) {
    if (!b || index < 0 || index >= (int32_t)b->engines.size() || !eig_threads)
        return fail(WNLQG_E_INVALID, "invalid arguments");
    const wnlqg_engine *g = b->engines[index];
    *eig_threads = g->eig_tpb;
    return WNLQG_OK;
}

}  // extern "C"
