/*
 * wnlqg.h -- C ABI of the MI355X k-mesh response engine (libwnlqg_hip.so).
 *
 * The engine replaces the reference's dense-mesh Integral k-loop for one model and one
 * task bundle: everything between the per-worker workspace preparation
 * (reference src/Runtime/Execution/IntegralDriver.jl:475-572) and finish!
 * (IntegralDriver.jl:909).  Per-lane partial tensors come back already scaled exactly as
 * consume_point! leaves them in state.worker_data[lane] (IntegralDriver.jl:820-874), so the
 * reference's _reduce_integral_states! (src/Runtime/Reduction/DeterministicReduction.jl:2-37)
 * and its writers run unchanged.
 *
 * Conventions
 *   - Complex values are interleaved (re, im) doubles.
 *   - Arrays described as "column-major [d1, d2, ...]" use Julia's layout (d1 fastest), so a
 *     Julia Array{ComplexF64} can be passed by pointer without copying.
 *   - All entry points return 0 on success or a negative WNLQG_E_* code; the message is
 *     available from wnlqg_last_error().  Nothing is thrown across the ABI.
 *   - One engine owns one HIP stream on the device that was current at wnlqg_create; calls
 *     on one engine must not overlap.  Calls block until their results are in host memory.
 */
#ifndef WNLQG_H
#define WNLQG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WNLQG_ABI_VERSION 9

/* Status codes. */
#define WNLQG_OK 0
#define WNLQG_E_INVALID (-1)     /* invalid argument or inconsistent sizes */
#define WNLQG_E_NONFINITE (-2)   /* non-finite model input or non-finite result */
#define WNLQG_E_GAP (-3)         /* spectral isolation violated (OrthogonalBandDerivatives.jl:74-78) */
#define WNLQG_E_HIP (-4)         /* HIP runtime error */
#define WNLQG_E_OOM (-5)         /* device allocation failed */
#define WNLQG_E_UNSUPPORTED (-6) /* configuration outside what this build implements */
#define WNLQG_E_NOT_CONVERGED (-7) /* the Jacobi eigensolver reached its sweep cap (40) at some k point without
                                      meeting its stopping test (LAPACK's zheevr reports INFO > 0 there) */

/* Response quantities (bitmask). Output states are ordered by increasing bit.  One engine serves one
 * reference driver family: SC and/or IC (IntegralDriver.jl), or BCD alone (BerryMultipoleDriver.jl),
 * or intrinsic AHC alone (SpectralResponseDriver.jl). */
#define WNLQG_SHIFT_CURRENT 1     /* Conventional SC, ConventionalShiftCurrent.jl:527-884 */
#define WNLQG_INJECTION_CURRENT 2 /* Conventional IC, ConventionalInjectionCurrent.jl:23-140 */
#define WNLQG_BERRY_CURVATURE_DIPOLE 4 /* BCD, Fermi-surface form: BerryMultipoleMoments.jl:7-45 */
#define WNLQG_INTRINSIC_HALL 8    /* intrinsic AHC, Matrix method: IntrinsicHallOccupations.jl:8-44 */
#define WNLQG_SECOND_HARMONIC 16  /* SHG chi / sigma(-2w; w, w): SecondHarmonicKernels.jl:115-294 */

/* Broadening (TransitionScreening.jl:1-29). */
#define WNLQG_GAUSSIAN 0
#define WNLQG_LORENTZIAN 1

/* Arithmetic forms. */
#define WNLQG_PRECISION_FP64 0  /* FP64 everywhere (plain GPU path) */
#define WNLQG_PRECISION_MIXED 1 /* SC / IC only.  Arbitrary k points and non-separable meshes: split-FP16
                                   (3 products) MFMA interpolation with FP64 partial sums.  Separable meshes
                                   (even n_w in [16, 64]): FP64 plane contraction whose channels after H are
                                   stored in FP32 (half the channel traffic).  Always: FP64 eigensolve,
                                   rotation arithmetic, denominators and vertices, Kahan-compensated k-sum */
#define WNLQG_PRECISION_MIXED_HIGH 2 /* split-FP16 with 4 products */
#define WNLQG_PRECISION_SPLIT_BF16 3 /* split-BF16 (3 products) on BF16 MFMA with FP32 accumulation and FP64
                                        partial sums; otherwise as MIXED.  Comparison form: the rounding of
                                        the coefficient operand is the same at every k, so its error floor
                                        (~1e-5 of the spectrum) does not average out over the mesh */

/* Interpolation backends for wnlqg_run_mesh (the role of fourier_backend, FourierExecutionPlan.jl:394-480). */
#define WNLQG_FOURIER_AUTO 0     /* separable plane contraction on meshes where it plans, else direct */
#define WNLQG_FOURIER_DIRECT 1   /* k x R contraction on MFMA (the only form for arbitrary k points) */
#define WNLQG_FOURIER_PLANES 2   /* separable: R3 sum per k3 plane, R2 sum per k2 line, R1 sum per k point */
#define WNLQG_FOURIER_MIXED 3    /* the reference's fourier_backend = "mixed" (NKdiv x NKFFT Mixed-FFT, the same
                                    uniform-mesh Fourier sums): the separable contraction, with the reference's
                                    errors: WNLQG_E_UNSUPPORTED from wnlqg_create when the contraction cannot
                                    plan the model, from wnlqg_run_mesh when it cannot plan the mesh, and from
                                    wnlqg_run_points (weighted k-point lists are adaptive / region nodes,
                                    which the reference evaluates with direct sums only,
                                    FourierExecutionPlan.jl:425-428) */

/*
 * Tight-binding model (TightBindingModel, reference src/Core/Models.jl:11-75; the fields of
 * read_wannier_tb, src/IO/WannierTB.jl:53-111).
 */
typedef struct wnlqg_model {
    int32_t num_orbitals;            /* n */
    int32_t num_r_vectors;           /* N_R */
    const double *lattice;           /* [9] Julia lattice[3,3] column-major; rows are a1,a2,a3 (Angstrom) */
    const int32_t *r_vectors;        /* [3*N_R] Julia r_vectors[3,N_R] column-major */
    const int32_t *r_degeneracies;   /* [N_R] */
    const double *hamiltonian_r;     /* complex, column-major [n,n,N_R] (eV) */
    const double *position_r;        /* complex, column-major [n,n,3,N_R] (Angstrom), or NULL for A = 0 */
} wnlqg_model;

/*
 * Resolved task parameters (subset of EffectiveTaskConfig, reference src/Runtime/Config.jl:4-170).
 * Wannier-center convention: Convention_II (the default, Config.jl:154) only.
 */
typedef struct wnlqg_task {
    int32_t quantities;               /* WNLQG_SHIFT_CURRENT | WNLQG_INJECTION_CURRENT, or one of
                                         WNLQG_BERRY_CURVATURE_DIPOLE, WNLQG_INTRINSIC_HALL */
    int32_t num_photon_energies;      /* n_w >= 1 (SC / IC; ignored otherwise) */
    const double *photon_energies;    /* [n_w] eV, ascending (SC / IC) */
    int32_t broadening_type;          /* WNLQG_GAUSSIAN or WNLQG_LORENTZIAN */
    double broadening;                /* eV (> 0) */
    double fermi_energy;              /* eV */
    double temperature;               /* K (>= 0) */
    double denominator_regularization;/* eta in eV (>= 0); 0 = bare 1/dE with isolation check */
    double transition_window_factor;  /* Lorentzian screening window, in broadenings */
    int32_t band_window_size;         /* -1 = all bands (only -1 or a window covering all bands) */
    int32_t precision;                /* WNLQG_PRECISION_* */
    int32_t batch_size;               /* k points per device batch; 0 = automatic */
    int32_t fourier_backend;          /* WNLQG_FOURIER_* (meshes only; points always use the direct contraction) */
    /* BCD / AHC (BerryMultipoleDCParameters, BerryMultipoleDC.jl:6-55; TaskConfiguration.jl:568-578) */
    int32_t num_fermi_energies;       /* n_mu, 1 .. 64 */
    const double *fermi_energies;     /* [n_mu] eV */
    double fermi_surface_broadening;  /* eta_fs_ev > 0: Gaussian -f'(E) at T = 0 (BCD); must be 0 when T > 0 */
    double degeneracy_threshold;      /* eV: BCD spectral isolation threshold, SHG one-band isolation
                                         (reference default 0.002) */
    /* SHG (Config.jl:107-109; SecondHarmonicDriver.jl:146-168).  SHG uses photon_energies, broadening,
     * broadening_type, fermi_energy, temperature and degeneracy_threshold as well. */
    int32_t num_shg_components;       /* 0 = all 18 in the reference order (a; bc = xx,yy,zz,yz,xz,xy) */
    const int32_t *shg_components;    /* [3 * num_shg_components] 0-based (a, b, c) */
    double shg_low_frequency_broadening;    /* eV (> 0; reference default 0.025) */
    double shg_intermediate_regularization; /* eV (> 0) */
    int32_t shg_eta_correction;       /* 1: eta-corrected optical position derivative (reference default) */
} wnlqg_task;

/* Counters of the last run (device event timers, milliseconds).  wnlqg_run_mesh runs batch b - 1's
 * screening, rotation, response and k-sums on a second stream beside batch b's interpolation and
 * eigensystem, so the stage timers span GPU time the two streams share and sum to more than total_ms;
 * with the environment variable WNLQG_MESH_OVERLAP=0 it runs one stream in program order and each
 * stage timer holds its kernels alone. */
typedef struct wnlqg_stats {
    double total_ms;        /* whole k-loop: phase, interpolation, eigen, screening, response, k-sum */
    double interp_ms;       /* interpolation GEMM kernel */
    double eig_ms;          /* Jacobi eigensolver kernel */
    double rotate_ms;       /* U^+ X U kernel (SC / IC: k_rotate_sc alone) */
    double pairs_ms;        /* screening + generalized derivative + response kernels */
    double accum_ms;        /* frequency contraction + deterministic k-sum */
    int64_t kpoints;        /* k points evaluated */
    int64_t active_pairs;   /* band pairs contributing */
    int32_t batches;        /* device batches */
    int32_t interp_launches;/* launches of the interpolation kernel */
    double interp_flops;    /* algorithmic interpolation flops: 8 * N_R * n^2 * C per k point */
    int32_t max_sweeps;     /* largest Jacobi sweep count over the k points */
    int32_t separable;      /* 1 if the mesh used the separable plane contraction */
    double interp_bytes;    /* separable contraction: interpolated channels written + coefficient box read per plane */
    double mean_sweeps;     /* Jacobi sweeps per k point, averaged over the mesh */
    double warm_ms;         /* warm-start transform U_a^+ H U_a of the mesh eigensolver */
    int32_t eig_launches;   /* eigensolver kernel launches (2 per batch when warm-started) */
    int32_t rotate_row_points;     /* SC / IC element points whose dH and A were formed as rows at the pairs'
                                      bands (exactly Hermitian models); a subset of rotate_element_points */
    int64_t rotate_element_points; /* SC / IC points whose dd H and dA were reduced to the pairs' elements */
    int64_t rotate_full_points;    /* SC / IC points with every channel rotated (points without pairs: neither) */
    int64_t rotate_element_bands;  /* sum over the element points of the number of bands in their pairs */
} wnlqg_stats;

typedef struct wnlqg_engine wnlqg_engine;

/* ABI version (WNLQG_ABI_VERSION). */
int32_t wnlqg_version(void);

/* Copy the message of the last failing call on this thread into buf (NUL-terminated).
 * Returns the full message length. */
int32_t wnlqg_last_error(char *buf, size_t len);

/* Upload the model, build the R-space channels on the device and allocate the k-batch
 * working set (replaces prepare_real_space!, HamiltonianMatrixElements.jl:180-199, and the
 * workspace setup of IntegralDriver.jl:212-446).  *out is NULL on failure. */
int32_t wnlqg_create(const wnlqg_model *model, const wnlqg_task *task, wnlqg_engine **out);

/* Release all device memory of the engine. NULL is accepted. */
int32_t wnlqg_destroy(wnlqg_engine *engine);

/* Number of output states (set bits of task.quantities) and interpolation channels stored per k point. */
int32_t wnlqg_num_states(const wnlqg_engine *engine);
int32_t wnlqg_num_channels(const wnlqg_engine *engine);

/*
 * Dense Gamma-centred mesh k_i = i/N_i, axis 1 fastest (integral_kpoint!,
 * src/Core/CalculationUtils.jl:280-292), integrated with the reference's deterministic lanes
 * (src/Runtime/Setup/BundleAssembly.jl:35-47): lane L covers k indices
 * floor(L*N_k/n_lanes) .. floor((L+1)*N_k/n_lanes)-1 (0-based), 1 <= n_lanes <= N_k
 * (the reference uses n_lanes = min(N_k, 64)).
 *
 * lane_out: complex [n_lanes][n_states][n_w*27]; each state block is Julia column-major
 * [n_w,3,3,3] and equals sum_{k in lane} (1/N_k) * value_k with the state's prefactor
 * (ConventionalCurrentIntegralExecution.jl:243-245 for SC, :296-298 for IC).
 * total_out (optional, may be NULL): complex [n_states][n_w*27], the lanes summed in lane order.
 *
 * BCD engines: each lane block is [n_mu,3,3,3] column-major (complex, imaginary parts 0) holding the
 * lane's sum of raw per-point moments D[mu,a,b,c] = sum_n -f'(E_n) Omega^n_ab dE_n/dk_c (what
 * BerryMultipoleDriver.jl:279-285 leaves in buffers[lane]); total_out = lanes summed / N_k / |det lattice|,
 * the integrated moment that berry_dipole_dc_conductivity (BerryMultipoleDC.jl:95-113) converts.  A level
 * pair closer than degeneracy_threshold at any k returns WNLQG_E_GAP.
 * AHC engines: each lane block is [n_mu,3,3] column-major, the lane's sum of per-point
 * (-e^2/hbar 1e-20) Omega_ab(mu) (SpectralResponseDriver.jl:388-400); total_out = lanes summed / N_k /
 * (|det lattice| 1e-30), sigma_ab in S/m (SpectralResponseDriver.jl:541-552).
 * SHG engines: each lane block is [n_w, n_comp, 7, 2] column-major, the lane's sum of raw per-point
 * densities (the buffers[lane] of SecondHarmonicDriver.jl:178-185; term 1..7, quantity chi, sigma);
 * total_out = lanes summed, quantity q times factors[q] / N_k with factors = (e/(eps0 V) 1e12,
 * e^2/(hbar V)), V = |det lattice| (SecondHarmonicDriver.jl:135-139, 236-242).
 * stats may be NULL.
 */
int32_t wnlqg_run_mesh(wnlqg_engine *engine, const int32_t mesh[3], int32_t n_lanes,
                       double *lane_out, double *total_out, wnlqg_stats *stats);

/*
 * Weighted point list for SC / IC, BCD and AHC engines: the k-loop over the evaluation items of a
 * symmetry-reduced mesh (src/Runtime/Execution/IntegralDriver.jl:822-831: kpoint_index =
 * representatives[i], kpoint_weight = multiplicities[i], scale = kpoint_weight / num_kpoints), or any
 * weighted node set (sub-cell refinements of a mesh).  kpoints: column-major [3,m] fractional;
 * weights[i]: the scale of item i (multiplicity / N_k for a reduced mesh).  Lanes partition the m items
 * as wnlqg_run_mesh partitions the mesh: lane L covers items floor(L*m/n_lanes) ..
 * floor((L+1)*m/n_lanes)-1, 1 <= n_lanes <= m.  SC / IC: lane_out complex [n_lanes][n_states][n_w*27],
 * each block sum_{i in lane} weights[i] * value_i with the state's prefactor; the tensor projection of
 * apply_response_symmetry! (ResponseSymmetryRuntime.jl:1121-1146) is the caller's.  BCD / AHC: lane
 * blocks as in wnlqg_run_mesh, each sum_{i in lane} weights[i] * value_i of the per-point values that
 * wnlqg_run_mesh's lanes sum (the weights take the place of the mesh's 1/N_k); total_out = the lanes
 * summed in lane order,
 * divided by the measure (|det L|, and 1e-30 for AHC), so weights 1/N_k over a whole mesh give
 * wnlqg_run_mesh's total.  total_out is optional.  SHG engines return WNLQG_E_INVALID.
 */
int32_t wnlqg_run_points(wnlqg_engine *engine, const double *kpoints, const double *weights, int64_t m,
                         int32_t n_lanes, double *lane_out, double *total_out);

/*
 * Symmetry-reduced dense mesh for SC / IC engines with the separable interpolation (fourier_backend
 * auto or planes): the same integral as wnlqg_run_points over the mesh points reps[0..m-1] (strictly
 * increasing linear indices of the Gamma-centred mesh, axis 1 fastest: the orbit representatives of
 * response_symmetry_kmesh_orbits, ResponseSymmetryRuntime.jl:463-496, 0-based) with weights[i]
 * (multiplicity / N_k), but interpolated by the plane contraction: its first two stages run for every
 * plane, its last stage and everything after it only for the representatives.  Lanes, lane_out
 * and total_out as in wnlqg_run_points; the tensor projection is the caller's.  WNLQG_E_INVALID for
 * other families or a mesh the separable interpolation does not plan.
 */
int32_t wnlqg_run_mesh_reduced(wnlqg_engine *engine, const int32_t mesh[3], const int64_t *reps,
                               const double *weights, int64_t m, int32_t n_lanes, double *lane_out,
                               double *total_out);

/*
 * Per-point values (point_value!, IntegralDriver.jl:596, with num_kpoints = 1) at m arbitrary
 * fractional k points.  kpoints: column-major [3,m].  out: complex [m][n_states][n_w*27]; for BCD
 * [m][n_mu*27] raw moments, for AHC [m][n_mu*9] scaled curvature, for SHG [m][n_w*n_comp*14] raw
 * densities (the per-point lane contributions).
 */
int32_t wnlqg_eval_points(wnlqg_engine *engine, const double *kpoints, int32_t m, double *out);

/*
 * wnlqg_eval_points that also returns the band energies of the same evaluation: energies [m][n],
 * ascending per point (eV).  No reference counterpart: the per-point spectrum is what the learned
 * refinement proposal reads as features of a cell, at no extra interpolation or eigensolve.
 */
int32_t wnlqg_eval_points_spectrum(wnlqg_engine *engine, const double *kpoints, int32_t m, double *out,
                                   double *energies);

/*
 * Wannier-gauge interpolated channels at m points (the Fourier sums of
 * HamiltonianMatrixElements.jl:42-58,427-462 and PositionMatrixElements.jl:239-409).
 * out: complex [m][C][n][n] with C = wnlqg_num_channels(); channel order:
 *   0 H, 1-3 dH_a (a = x,y,z), 4-9 ddH_ab for ab = xx,yy,zz,yz,xz,xy,
 *   10-12 A_a, 13-21 dA_ab (= Fourier of i R_b A_a, index 13 + 3a + b)  (the A channels
 *   exist only when the model has a position operator and a quantity needs them).
 *   BCD / AHC engines store no ddH: 0 H, 1-3 dH_a, 4-6 A_a, 7-15 dA_ab (index 7 + 3a + b).
 * Element [i][j] is row i, column j.
 */
int32_t wnlqg_fourier_points(wnlqg_engine *engine, const double *kpoints, int32_t m, double *out);

/*
 * Eigensystem of the upper triangle of H(k) (DenseSpectra.hermitian_spectrum,
 * src/Core/ESKernel/DenseSpectra.jl:14-16) with the canonical column phase
 * (canonicalize_eigenvectors!, HamiltonianMatrixElements.jl:6-25).
 * energies: [m][n] ascending.  vectors: complex [m][n][n], element [i][j] = component i of
 * eigenvector j.
 */
int32_t wnlqg_spectrum_points(wnlqg_engine *engine, const double *kpoints, int32_t m,
                              double *energies, double *vectors);

/* Block until all work of the engine's stream is complete. */
int32_t wnlqg_synchronize(wnlqg_engine *engine);

/*
 * Batching form: `count` independent (model, task, mesh) integrals of uneven n_w run by one
 * hipGraph launch (the reference runs such bundles one task at a time through the same
 * execute_prepared_tasks! seam, IntegralDriver.jl:1079 / BerryMultipoleDriver.jl / SpectralResponseDriver.jl).
 * wnlqg_batch_create builds one engine per entry (tasks with batch_size 0 share a working-set
 * budget), plans every mesh, optionally autotunes the eigensolver's threads per k point for each n_w
 * (autotune_tiles != 0), and captures all kernel sequences into one graph whose entries run on four
 * concurrent streams.  meshes: [3*count]; n_lanes: [count].
 * wnlqg_batch_run replays the graph and writes, for entry i, exactly what wnlqg_run_mesh writes into
 * lane_out[i] / total_out[i] (total_out or any total_out[i] may be NULL); elapsed_ms = device time of the graph.
 */
typedef struct wnlqg_batch wnlqg_batch;
int32_t wnlqg_batch_create(int32_t count, const wnlqg_model *models, const wnlqg_task *tasks, const int32_t *meshes,
                           const int32_t *n_lanes, int32_t autotune_tiles, wnlqg_batch **out);
int32_t wnlqg_batch_run(wnlqg_batch *batch, double *const *lane_out, double *const *total_out, double *elapsed_ms);
/* the eigensolver form chosen for entry `index`: 64, 128 or 256 (threads per k point of the LDS Jacobi),
 * 1 (register-resident eigenvectors, n_w <= 32), 2 (deferred eigenvectors: H-only sweeps with recorded
 * rotations applied to V afterwards, n_w <= 32) or 0 (size-class default) */
int32_t wnlqg_batch_tuning(const wnlqg_batch *batch, int32_t index, int32_t *eig_threads);
int32_t wnlqg_batch_destroy(wnlqg_batch *batch);

/*
 * Learned refinement proposal.  No counterpart in the reference (it has no learned component); for
 * callers that rank the cells of a k mesh for adaptive refinement from per-cell features of the base
 * mesh, trained on the refinement gains of cheaper coarse runs.  The integrand itself is always
 * computed exactly by the engine above.
 *
 * An 8 -> 64 -> 64 -> 1 multilayer perceptron with SiLU activations in FP32.  Parameter vector:
 * W1[64][8] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3 (row-major, WNLQG_MLP_PARAMS floats).
 * Samples x are [n][8] floats, targets and outputs [n].  The loss is the mean squared error; the
 * gradient is deterministic (fixed per-wave sample ranges summed in wave order).  Training is Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of a per-epoch shuffle seeded by `seed`.
 * Errors: the same WNLQG_E_* codes, message from wnlqg_mlp_last_error.
 */
#define WNLQG_MLP_INPUTS 8
#define WNLQG_MLP_HIDDEN 64
#define WNLQG_MLP_PARAMS 4801

typedef struct wnlqg_mlp wnlqg_mlp;

int32_t wnlqg_mlp_create(uint64_t seed, wnlqg_mlp **out);   /* Glorot-uniform weights, zero biases */
int32_t wnlqg_mlp_destroy(wnlqg_mlp *mlp);
int32_t wnlqg_mlp_last_error(char *buf, size_t len);
int32_t wnlqg_mlp_get_params(wnlqg_mlp *mlp, float *params);
int32_t wnlqg_mlp_set_params(wnlqg_mlp *mlp, const float *params);  /* also resets the Adam moments */
int32_t wnlqg_mlp_predict(wnlqg_mlp *mlp, const float *x, int64_t n, float *y);
/* gradient of the mean squared error over all n samples, and the loss */
int32_t wnlqg_mlp_gradient(wnlqg_mlp *mlp, const float *x, const float *t, int64_t n, float *grad, float *loss);
/*
 * wnlqg_mlp_gradient without any allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  x [n][8], t [n] and grad [WNLQG_MLP_PARAMS + 1] are DEVICE pointers; grad receives the
 * gradient of the mean squared error followed by the SUM of squared errors (divide by n for the loss).
 * The per-wave partial gradients live in a library-owned static device array of
 * WNLQG_MLP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of buffers of the mlp object;
 * the call makes no hipMalloc / hipFree.  1 <= n <= WNLQG_MLP_NOALLOC_MAX_SAMPLES, else WNLQG_E_INVALID.
 * stream: a hipStream_t to enqueue on and return immediately, or NULL for the mlp's own stream, synchronized
 * before returning.  Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU
 * (other streams, other threads, other mlp objects) race and must be serialized by the caller.
 */
#define WNLQG_MLP_NOALLOC_SCRATCH_BYTES (512ULL * 4802ULL * 4ULL)
#define WNLQG_MLP_NOALLOC_MAX_SAMPLES (1LL << 40)
int32_t wnlqg_mlp_gradient_noalloc(wnlqg_mlp *mlp, const float *x, const float *t, int64_t n, float *grad,
                                   void *stream);
/* `steps` Adam steps of `batch` samples; loss_history[steps] (may be NULL) receives each batch loss */
int32_t wnlqg_mlp_train(wnlqg_mlp *mlp, const float *x, const float *t, int64_t n, int32_t steps, int32_t batch,
                        float lr, uint64_t seed, float *loss_history);

#ifdef __cplusplus
}
#endif

#endif /* WNLQG_H */
