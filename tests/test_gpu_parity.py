"""GPU parity of the HIP engine against the FP64 oracle (oracle/kloop.py), through the C ABI.

Bar: FP64 path within 1e-10 relative (L2 over the omega spectrum, per component) of the oracle;
integer bookkeeping (lanes) exact.  Each test prints the case it runs before the GPU call."""
import os
import sys

import numpy as np
import pytest

from oracle import kloop
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import TightBindingModel, config_model, read_wannier_tb, synthetic_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def mk_task(**kw):
    base = dict(photon_energies=np.linspace(0.0, 4.0, 200), broadening=0.02, shift_current=True,
                injection_current=True)
    base.update(kw)
    return base


def engine(model, **kw):
    return E.Engine(model, E.EngineTask(**mk_task(**kw)))


def oracle_task(**kw):
    t = mk_task(**kw)
    return kloop.Task(photon_energies=t["photon_energies"], broadening=t["broadening"],
                      broadening_type=t.get("broadening_type", "Gaussian"), fermi_energy=t.get("fermi_energy", 0.0),
                      temperature=t.get("temperature", 0.0),
                      denominator_regularization=t.get("denominator_regularization", 1e-3),
                      transition_window_factor=t.get("transition_window_factor", 5.0),
                      shift_current=t["shift_current"], injection_current=t["injection_current"])


MODELS = {
    "fixture_n4": lambda: read_wannier_tb(FIXTURE),
    "cfg0_n8": lambda: config_model(0),
    "n12_pos": lambda: synthetic_model(12, 2, 7, position_scale=0.1),
    "n32_pos": lambda: synthetic_model(32, 2, 11, position_scale=0.1),
}


def channels_oracle(model, kp):
    w = kloop.interpolate(model, kp)
    n = model.num_orbitals
    sym = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    chans = [w["H"][:, None]] + [w["dH"][:, a][:, None] for a in range(3)]
    chans += [w["ddH"][:, a, b][:, None] for a, b in sym]
    if model.position_r is not None:
        chans += [w["A"][:, a][:, None] for a in range(3)]
        chans += [w["dA"][:, a, b][:, None] for a in range(3) for b in range(3)]
    return np.concatenate(chans, axis=1)


@pytest.mark.parametrize("name", list(MODELS))
def test_fourier_channels(name):
    model = MODELS[name]()
    rng = np.random.default_rng(5)
    kp = rng.uniform(-1, 1, (37, 3))
    say(f"[fourier] model={name} n={model.num_orbitals} nr={model.num_r_vectors} points=37")
    with engine(model) as eng:
        got = eng.fourier_points(kp)
    exp = channels_oracle(model, kp)
    assert got.shape == exp.shape
    scale = np.max(np.abs(exp), axis=(2, 3), keepdims=True) + 1e-300
    err = np.max(np.abs(got - exp) / scale)
    say(f"  max rel err {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("name", list(MODELS))
def test_spectrum(name):
    model = MODELS[name]()
    rng = np.random.default_rng(6)
    kp = rng.uniform(0, 1, (41, 3))
    say(f"[spectrum] model={name} points=41")
    with engine(model) as eng:
        e, u = eng.spectrum_points(kp)
    hk = kloop.interpolate(model, kp)["H"]
    e0, u0 = kloop.hermitian_spectrum(hk)
    scale = np.max(np.abs(e0))
    assert np.max(np.abs(e - e0)) < 1e-12 * scale
    # residual of the upper-triangle Hermitian matrix and orthonormality
    hu = np.triu(hk) + np.conj(np.swapaxes(np.triu(hk, 1), 1, 2))
    hu[:, np.arange(model.num_orbitals), np.arange(model.num_orbitals)] = np.real(np.diagonal(hk, axis1=1, axis2=2))
    res = hu @ u - u * e[:, None, :]
    assert np.max(np.abs(res)) < 1e-11 * scale
    eye = np.conj(np.swapaxes(u, 1, 2)) @ u
    assert np.max(np.abs(eye - np.eye(model.num_orbitals))) < 1e-12
    # canonical phase: first largest-|.| entry of each column is real and >= 0
    piv = np.argmax(np.abs(u0), axis=1)
    pv = np.take_along_axis(u, piv[:, None, :], axis=1)[:, 0, :]
    assert np.max(np.abs(pv.imag)) < 1e-12 and np.min(pv.real) >= 0
    gaps = np.min(np.diff(e0, axis=1), axis=1)
    iso = gaps > 1e-6
    assert np.max(np.abs(u[iso] - u0[iso])) < 1e-8


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("smearing", ["Gaussian", "Lorentzian"])
def test_point_values(name, smearing):
    model = MODELS[name]()
    rng = np.random.default_rng(7)
    kp = rng.uniform(0, 1, (23, 3))
    kw = dict(broadening_type=smearing, photon_energies=np.linspace(0, 6, 150), broadening=0.05,
              fermi_energy=0.1 if name != "fixture_n4" else 0.0)
    say(f"[points] model={name} smearing={smearing}")
    with engine(model, **kw) as eng:
        got = eng.eval_points(kp)
    exp = kloop.point_values(model, oracle_task(**kw), kp)
    for key in ("sc", "ic"):
        r = rel(got[key], exp[key])
        say(f"  {key} rel {r:.3e} norm {np.linalg.norm(exp[key]):.3e}")
        assert r < 1e-10


def test_point_values_finite_temperature_and_eta():
    model = MODELS["n12_pos"]()
    kp = np.random.default_rng(8).uniform(0, 1, (19, 3))
    kw = dict(temperature=300.0, denominator_regularization=0.02, photon_energies=np.linspace(0.2, 5, 90))
    say("[points] finite T, eta = 0.02")
    with engine(model, **kw) as eng:
        got = eng.eval_points(kp)
    exp = kloop.point_values(model, oracle_task(**kw), kp)
    for key in ("sc", "ic"):
        assert rel(got[key], exp[key]) < 1e-10


def test_bare_denominators_and_gap_error():
    """denominator_regularization = 0 selects the bare 1/dE (OrthogonalBandDerivatives.jl:66-82): on a model
    with isolated levels it matches the oracle; on the direct sum of a model with itself every level is
    exactly doubly degenerate, and the screening kernel's isolation check returns WNLQG_E_GAP."""
    model = MODELS["n12_pos"]()
    kp = np.random.default_rng(9).uniform(0, 1, (17, 3))
    kw = dict(denominator_regularization=0.0, photon_energies=np.linspace(0.0, 5.0, 80), broadening=0.05)
    say("[points] eta = 0, n12")
    with engine(model, **kw) as eng:
        got = eng.eval_points(kp)
    exp = kloop.point_values(model, oracle_task(**kw), kp)
    for key in ("sc", "ic"):
        r = rel(got[key], exp[key])
        say(f"  {key} rel {r:.3e}")
        assert r < 1e-10
    n = model.num_orbitals
    h = np.zeros((model.num_r_vectors, 2 * n, 2 * n), dtype=np.complex128)
    a = np.zeros((model.num_r_vectors, 3, 2 * n, 2 * n), dtype=np.complex128)
    h[:, :n, :n] = h[:, n:, n:] = model.hamiltonian_r
    a[:, :, :n, :n] = a[:, :, n:, n:] = model.position_r
    doubled = TightBindingModel(model.lattice, model.r_vectors, model.r_degeneracies, h, a, name="doubled")
    say("[points] eta = 0, doubly degenerate n24")
    with engine(doubled, **kw) as eng:
        with pytest.raises(E.WnlqgError) as err:
            eng.eval_points(kp[:4])
    assert err.value.code == -3


@pytest.mark.parametrize("name,mesh", [("cfg0_n8", (8, 8, 8)), ("fixture_n4", (5, 3, 2)), ("n32_pos", (4, 3, 3)),
                                       ("n12_pos", (1, 1, 1))])
def test_mesh_integral(name, mesh):
    """Default backend (separable plane contraction on meshes) against the oracle."""
    model = MODELS[name]()
    kw = dict(photon_energies=np.linspace(0, 5, 120), broadening=0.04)
    say(f"[mesh] model={name} mesh={mesh}")
    with engine(model, **kw) as eng:
        lanes, total, stats = eng.run_mesh(mesh)
        lanes2, total2, _ = eng.run_mesh(mesh)
    ol, ot = kloop.integrate_mesh(model, oracle_task(**kw), mesh, batch=64)
    for key in ("sc", "ic"):
        assert lanes[key].shape == ol[key].shape
        r = kloop.spectrum_relative_error(total[key], ot[key])
        say(f"  {key} worst component rel {r.max():.3e}")
        assert r.max() < 1e-10
        assert rel(lanes[key], ol[key]) < 1e-10
        # determinism: repeated runs are bitwise identical
        assert np.array_equal(lanes[key], lanes2[key])
        # total = lanes summed in lane order
        acc = np.zeros_like(total[key])
        for l in range(lanes[key].shape[0]):
            acc = acc + lanes[key][l]
        assert np.array_equal(acc, total[key])
    assert stats["kpoints"] == int(np.prod(mesh))


@pytest.mark.parametrize("omega_max,paths", [(6.0, "elements"), (7.0, "both"), (24.0, "full")])
def test_element_rotation_paths(omega_max, paths):
    """The SC/IC rotation reduces dd H and dA to the active pairs' elements for points with at most 8 pairs
    and rotates every channel otherwise.  On the benchmark model (bands over +-30 eV) photon energies up to
    6 eV give 2-7 pairs per point (element path only), up to 7 eV a mix, up to 24 eV 70-95 (full path only);
    each case against the oracle at the FP64 bar, with the device counts of both paths."""
    model = config_model(1)
    kw = dict(photon_energies=np.linspace(0, omega_max, 200), broadening=0.01)
    mesh = (4, 3, 3)
    say(f"[element-rotation] benchmark model, omega <= {omega_max} eV, mesh {mesh}")
    with engine(model, **kw) as eng:
        _, total, stats = eng.run_mesh(mesh)
    _, ot = kloop.integrate_mesh(model, oracle_task(**kw), mesh, batch=36)
    for key in ("sc", "ic"):
        r = kloop.spectrum_relative_error(total[key], ot[key]).max()
        say(f"  {key} worst component rel {r:.3e}")
        assert r < 1e-10
    el, full = stats["rotate_element_points"], stats["rotate_full_points"]
    say(f"  element points {el}, full points {full}, bands {stats['rotate_element_bands']}")
    assert el + full <= int(np.prod(mesh))
    assert (el > 0) == (paths in ("elements", "both")) and (full > 0) == (paths in ("full", "both"))


@pytest.mark.parametrize("name,omega_max,prec", [("cfg1", 6.0, "fp64"), ("cfg1", 7.0, "fp64"), ("n24", 5.0, "fp64"),
                                                 ("fixture", 4.0, "fp64"), ("cfg1", 6.0, "mixed"),
                                                 ("cfg1", 7.0, "mixed"), ("n24", 5.0, "mixed")])
def test_hermitian_row_rotation(monkeypatch, name, omega_max, prec):
    """For a model that is exactly Hermitian (H(-R) = H(R)^+, A(-R) = A(R)^+; the synthetic models are built
    that way) the SC/IC element rotation forms the velocity and position channels as rows at the pairs' bands
    and their conjugates as columns.  Against the full rotation of the same engine settings
    (WNLQG_ROTATE_HERMITIAN=0) the mesh integrals agree to 1e-12 and both match the oracle; the reference's
    fixture model (H(-R) != H(R)^+) keeps the full rotation, so both settings agree bitwise there.  In the
    mixed form both settings read the same FP32 channels and form FP64 products (the rows path with 4-row
    MFMA tiles up to 8 bands, 16-row tiles above), so they agree to 1e-12 as well; against the FP64 oracle the
    FP32 channel storage bounds the error at 1e-6."""
    model = {"cfg1": lambda: config_model(1), "n24": lambda: synthetic_model(24, 2, 77, position_scale=0.1),
             "fixture": lambda: read_wannier_tb(FIXTURE)}[name]()
    kw = dict(photon_energies=np.linspace(0, omega_max, 120), broadening=0.02)
    mesh = (4, 3, 3)
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("WNLQG_ROTATE_HERMITIAN", flag)
        say(f"[herm-rotation] {name} omega <= {omega_max} eV, {prec}, rows-only allowed: {flag}")
        with engine(model, precision=prec, **kw) as eng:
            lanes, total, stats = eng.run_mesh(mesh)
        res[flag] = (lanes, total)
    monkeypatch.delenv("WNLQG_ROTATE_HERMITIAN")
    _, ot = kloop.integrate_mesh(model, oracle_task(**kw), mesh, batch=36)
    for key in ("sc", "ic"):
        d = rel(res["1"][0][key], res["0"][0][key])
        say(f"  {key}: rows-only vs full {d:.2e}")
        assert d < 1e-12
        if name == "fixture":
            assert np.array_equal(res["1"][0][key], res["0"][0][key])
        for flag in ("1", "0"):
            assert kloop.spectrum_relative_error(res[flag][1][key], ot[key]).max() < (1e-10 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("n,mesh", [(32, (5, 3, 1)), (16, (7, 1, 1)), (20, (5, 1, 2)), (32, (3, 1, 1)),
                                    (24, (9, 2, 1)), (32, (48, 1, 1))])
def test_warm_started_eigensolver(n, mesh):
    """Mesh eigensolver warm start (groups of 4 consecutive k points; the anchor solved from the
    identity, the others from the anchor's eigenvectors): short last groups (15, 7, 5 and 9 points per
    row), padded sizes (n = 20, 24) and a row too short for a group (3 points, cold solve) all match the
    oracle.  On the fine row (48 points) the warm-started points need fewer sweeps than the anchors."""
    model = synthetic_model(n, 2, 40 + n, position_scale=0.1)
    kw = dict(photon_energies=np.linspace(0, 5, 80), broadening=0.04)
    say(f"[warm-eig] n={n} mesh={mesh}")
    with engine(model, **kw) as eng:
        lanes, total, st = eng.run_mesh(mesh)
    ol, ot = kloop.integrate_mesh(model, oracle_task(**kw), mesh, batch=64)
    for key in ("sc", "ic"):
        r = kloop.spectrum_relative_error(total[key], ot[key])
        say(f"  {key} worst component rel {r.max():.3e}")
        assert r.max() < 1e-10
        assert rel(lanes[key], ol[key]) < 1e-10
    warm = mesh[0] * mesh[1] >= 4
    say(f"  sweeps mean {st['mean_sweeps']:.2f} max {st['max_sweeps']} eig launches {st['eig_launches']}")
    # these small meshes fit one batch (planes grouped): every batch in groups of 4
    assert st["batches"] == 1
    assert st["eig_launches"] == (2 if warm else 1)
    assert 0 < st["mean_sweeps"] <= st["max_sweeps"]
    if mesh[0] >= 48:
        assert st["mean_sweeps"] < st["max_sweeps"] - 0.5


@pytest.mark.parametrize("n,mesh", [(32, (10, 10, 5)), (16, (12, 8, 6)), (48, (9, 9, 4))])
def test_plane_warm_started_eigensolver(n, mesh):
    """One plane per batch (planes of 64-128 points with 128-point batches): planes after the first start
    every point from the eigenvectors of the same point one k3 plane earlier (the other pipeline buffer):
    parity with the oracle, one eigensolver launch per later plane."""
    model = synthetic_model(n, 2, 70 + n, position_scale=0.1)
    kw = dict(photon_energies=np.linspace(0, 5, 80), broadening=0.04, batch_size=128)
    say(f"[plane-warm] n={n} mesh={mesh}")
    with engine(model, **kw) as eng:
        lanes, total, st = eng.run_mesh(mesh)
        lanes2, _, _ = eng.run_mesh(mesh)
    okw = {k: v for k, v in kw.items() if k != "batch_size"}
    ol, ot = kloop.integrate_mesh(model, oracle_task(**okw), mesh, batch=64)
    for key in ("sc", "ic"):
        r = kloop.spectrum_relative_error(total[key], ot[key])
        say(f"  {key} worst component rel {r.max():.3e}")
        assert r.max() < 1e-10
        assert rel(lanes[key], ol[key]) < 1e-10
        assert np.array_equal(lanes[key], lanes2[key])
    say(f"  sweeps mean {st['mean_sweeps']:.2f} max {st['max_sweeps']} eig launches {st['eig_launches']}")
    assert st["batches"] == mesh[2] and st["eig_launches"] == 2 + (mesh[2] - 1)
    assert st["mean_sweeps"] <= st["max_sweeps"]


@pytest.mark.parametrize("n", [8, 12, 16, 24, 32])
def test_deferred_eigensolver_matches_register_v(monkeypatch, n):
    """The deferred-eigenvector Jacobi (H-only sweeps recording each round's rotations, then the recorded
    rounds applied to register rows of V) performs the register-V kernel's rotations in the same order.  Cold
    solves (point spectra), group and previous-plane warm starts (mesh runs) and the continuation kernel
    (record depth lowered to 1 and 2 sweeps, so that most points are finished from their saved state) give
    the register-V kernel's eigenpairs and mesh integrals to within 1e-13 (bitwise in practice) and the same
    sweep counts; the mesh integrals also match the oracle."""
    model = synthetic_model(n, 2, 90 + n, position_scale=0.1)
    kp = np.random.default_rng(n).uniform(-0.5, 0.5, (300, 3))
    kw = dict(photon_energies=np.linspace(0, 5, 60), broadening=0.04, batch_size=128)
    meshes = [(9, 7, 3), (10, 10, 5)]
    runs = {}
    for form, rec in (("regv", None), ("defer", None), ("defer", "2"), ("defer", "1")):
        monkeypatch.setenv("WNLQG_EIG_KERNEL", form)
        if rec:
            monkeypatch.setenv("WNLQG_EIGD_RECORD_SWEEPS", rec)
        else:
            monkeypatch.delenv("WNLQG_EIGD_RECORD_SWEEPS", raising=False)
        say(f"[eig-defer] n={n} kernel={form} record_sweeps={rec or 8}")
        with engine(model, **kw) as eng:
            e, u = eng.spectrum_points(kp)
            mesh = [eng.run_mesh(m) for m in meshes]
        runs[(form, rec)] = (e, u, mesh)
    monkeypatch.delenv("WNLQG_EIG_KERNEL")
    monkeypatch.delenv("WNLQG_EIGD_RECORD_SWEEPS", raising=False)
    e0, u0, mesh0 = runs[("regv", None)]
    for key, (e, u, mesh) in runs.items():
        if key[0] == "regv":
            continue
        de = np.max(np.abs(e - e0)) / np.max(np.abs(e0))
        du = np.max(np.abs(u - u0))
        say(f"  {key}: eigenvalues {de:.2e}, eigenvectors {du:.2e}, bitwise {np.array_equal(e, e0) and np.array_equal(u, u0)}")
        assert de <= 1e-13 and du <= 1e-12
        for (l1, t1, s1), (l0, t0, s0) in zip(mesh, mesh0):
            assert s1["mean_sweeps"] == s0["mean_sweeps"] and s1["max_sweeps"] == s0["max_sweeps"]
            for q in ("sc", "ic"):
                assert rel(l1[q], l0[q]) < 1e-12
    okw = {k: v for k, v in kw.items() if k != "batch_size"}
    for m, (lanes, total, st) in zip(meshes, runs[("defer", None)][2]):
        ol, ot = kloop.integrate_mesh(model, oracle_task(**okw), m, batch=64)
        for q in ("sc", "ic"):
            assert kloop.spectrum_relative_error(total[q], ot[q]).max() < 1e-10
            assert rel(lanes[q], ol[q]) < 1e-10
        say(f"  mesh {m}: sweeps mean {st['mean_sweeps']:.2f} max {st['max_sweeps']}, oracle parity ok")


@pytest.mark.parametrize("backend,batch,mesh,nbatch", [("direct", 128, (6, 5, 7), 2), ("planes", 128, (6, 5, 7), 2),
                                                      ("planes", 128, (40, 6, 3), 6)])
def test_mesh_lane_counts_and_batches(backend, batch, mesh, nbatch):
    """Results do not depend on the lane count beyond rounding; small batches exercise the
    multi-batch chunk plan (lanes straddling batches, several planes per batch (4 + 3 planes of 30
    points), k2 sub-ranges of a plane)."""
    model = MODELS["cfg0_n8"]()
    kw = dict(photon_energies=np.linspace(0, 5, 60), broadening=0.04, batch_size=batch, fourier_backend=backend)
    say(f"[mesh] cfg0 {mesh} backend {backend} batch {batch}, lanes 64 vs 7")
    with engine(model, **kw) as eng:
        l64, t64, st = eng.run_mesh(mesh)
        l7, t7, _ = eng.run_mesh(mesh, n_lanes=7)
    assert st["batches"] == nbatch
    assert st["separable"] == (backend == "planes")
    okw = {k: v for k, v in kw.items() if k not in ("batch_size", "fourier_backend")}
    ol, ot = kloop.integrate_mesh(model, oracle_task(**okw), mesh, batch=50)
    for key in ("sc", "ic"):
        assert rel(t64[key], ot[key]) < 1e-10
        assert rel(t7[key], ot[key]) < 1e-10
        assert rel(l64[key], ol[key]) < 1e-10


@pytest.mark.parametrize("prec,tol", [("mixed", 2e-4), ("mixed6", 2e-6), ("bf16x3", 1e-3)])
def test_mixed_fourier_channels(prec, tol):
    """Split-FP16 (mixed, mixed6) and split-BF16 (bf16x3) interpolation: per-channel error relative to
    the channel's largest entry."""
    model = MODELS["n32_pos"]()
    kp = np.random.default_rng(5).uniform(-1, 1, (64, 3))
    say(f"[fourier-{prec}] n32")
    with engine(model, precision=prec) as eng:
        got = eng.fourier_points(kp)
    exp = channels_oracle(model, kp)
    scale = np.max(np.abs(exp), axis=(2, 3), keepdims=True) + 1e-300
    err = np.max(np.abs(got - exp) / scale)
    say(f"  max rel err {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-6), ("mixed6", 1e-6), ("bf16x3", 1e-4)])
def test_mixed_mesh_vs_fp64(prec, tol):
    """The mixed forms change only the direct k x R interpolation GEMM (meshes default to the separable
    FP64 contraction), so both engines run the direct backend."""
    model = MODELS["n12_pos"]()
    kw = dict(photon_energies=np.linspace(0, 5, 100), broadening=0.05, fourier_backend="direct")
    say(f"[mesh-{prec}] n12 8x8x8 direct")
    with engine(model, **kw) as e64:
        _, t64, s64 = e64.run_mesh((8, 8, 8))
    with engine(model, precision=prec, **kw) as emx:
        _, tmx, smx = emx.run_mesh((8, 8, 8))
    assert s64["separable"] == 0 and smx["separable"] == 0
    worst = 0.0
    for key in ("sc", "ic"):
        r = kloop.spectrum_relative_error(tmx[key], t64[key]).max()
        say(f"  {key} worst component rel {r:.3e}")
        worst = max(worst, r)
        assert r < tol
    assert worst > 0.0      # the reduced-precision GEMM did run


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-6), ("mixed6", 1e-11)])
def test_mixed_separable_mesh_vs_fp64(prec, tol):
    """On the separable mesh path the mixed form stores the interpolated channels after H in FP32 (the
    contraction computes in FP64; the rotation, eigensolver, vertices and compensated lane sums are FP64):
    within the metric's 1e-6 relative spectrum error of FP64, and visibly not FP64.  mixed6 keeps FP64
    channels and agrees to rounding."""
    model = MODELS["n32_pos"]()
    kw = dict(photon_energies=np.linspace(0, 6, 120), broadening=0.03)
    say(f"[mesh-sep-{prec}] n32 6x6x6 planes")
    with engine(model, **kw) as e64:
        _, t64, s64 = e64.run_mesh((6, 6, 6))
    with engine(model, precision=prec, **kw) as emx:
        _, tmx, smx = emx.run_mesh((6, 6, 6))
    assert s64["separable"] == 1 and smx["separable"] == 1
    worst = 0.0
    for key in ("sc", "ic"):
        r = kloop.spectrum_relative_error(tmx[key], t64[key]).max()
        say(f"  {key} worst component rel {r:.3e}")
        worst = max(worst, r)
        assert r < tol
    if prec == "mixed":
        assert worst > 1e-12      # the FP32 channels did run


@pytest.mark.parametrize("name", ["fixture_n4", "n32_pos"])
def test_planes_equal_direct(name):
    """The separable plane contraction reorders the same FP64 sum: channels agree to rounding."""
    model = MODELS[name]()
    kw = dict(photon_energies=np.linspace(0, 5, 80), broadening=0.04)
    say(f"[planes-vs-direct] {name}")
    with engine(model, fourier_backend="direct", **kw) as ed:
        _, td, sd = ed.run_mesh((6, 4, 5))
    with engine(model, fourier_backend="planes", **kw) as ep:
        _, tp, sp = ep.run_mesh((6, 4, 5))
    assert sd["separable"] == 0 and sp["separable"] == 1
    for key in ("sc", "ic"):
        assert rel(tp[key], td[key]) < 1e-11


def test_mixed_backend_runs_the_separable_contraction():
    """fourier_backend = "mixed" (the reference's Mixed-FFT name) runs the separable contraction on meshes,
    equal to the direct sums to rounding, and keeps the reference's errors: a weighted k-point list (adaptive /
    region nodes, FourierExecutionPlan.jl:425-428) and a mesh the contraction cannot plan are rejected with
    WNLQG_E_UNSUPPORTED."""
    model = MODELS["n32_pos"]()
    kw = dict(photon_energies=np.linspace(0, 5, 80), broadening=0.04)
    say("[mixed-backend] n32_pos 6x4x5 against direct")
    with engine(model, fourier_backend="direct", **kw) as ed:
        _, td, _ = ed.run_mesh((6, 4, 5))
    with engine(model, fourier_backend="mixed", batch_size=128, **kw) as em:
        _, tm, sm = em.run_mesh((6, 4, 5))
        assert sm["separable"] == 1
        for key in ("sc", "ic"):
            assert rel(tm[key], td[key]) < 1e-11
        with pytest.raises(E.WnlqgError) as ei:
            em.run_points(np.array([[0.1, 0.2, 0.3]]), np.array([1.0]))
        assert ei.value.code == -6
        with pytest.raises(E.WnlqgError) as ei:
            em.run_mesh((200, 1, 1))          # axis 0 longer than the 128-point batch
        assert ei.value.code == -6


@pytest.mark.parametrize("with_positions", [False, True])
def test_point_values_equal_projector_formulation(with_positions):
    """Per-point SC and IC of the engine (eta = 0) against the projector / band-velocity formulation of the
    theory notes in SI units (tests/test_oracle_response_pins.py), with no oracle in between; with
    positions the model carries A(0) = diag(tau) and the reference side uses the Convention-I matrix."""
    import test_oracle_response_pins as pins
    model, tau = pins._tau_model() if with_positions else (pins._sc_model(), None)
    kappa = pins.KAPPAS[0]
    omega, width, sigma, eta, _ = pins.si_point_values(model, kappa, tau=tau)
    say(f"projector pin on the GPU: positions={with_positions}")
    with E.Engine(model, E.EngineTask(photon_energies=omega, broadening=width, shift_current=True,
                                      injection_current=True, denominator_regularization=0.0)) as eng:
        vals = eng.eval_points(pins._frac(model, kappa)[None])
    assert np.max(np.abs(vals["sc"][0] - sigma)) <= 1e-6 * np.max(np.abs(sigma))
    assert np.max(np.abs(vals["ic"][0] - eta)) <= 1e-6 * np.max(np.abs(eta))


def test_eigensolver_sweep_cap_reports_not_converged(monkeypatch):
    """A sweep cap of 1 (set from the host before the engine is created) leaves every cold solve
    unconverged: the point and mesh entry points fail with WNLQG_E_NOT_CONVERGED instead of returning the
    unconverged eigenpairs; without the cap the same engine settings succeed."""
    model = config_model(0)
    kp = np.random.default_rng(5).uniform(0, 1, (16, 3))
    say("sweep cap 1: spectrum, eval and mesh must report non-convergence")
    monkeypatch.setenv("WNLQG_JACOBI_MAX_SWEEPS", "1")
    with engine(model) as eng:
        for call in (lambda: eng.spectrum_points(kp), lambda: eng.eval_points(kp), lambda: eng.run_mesh((4, 4, 4))):
            with pytest.raises(E.WnlqgError) as ei:
                call()
            assert ei.value.code == -7
    monkeypatch.delenv("WNLQG_JACOBI_MAX_SWEEPS")
    with engine(model) as eng:
        eng.eval_points(kp)
        eng.run_mesh((4, 4, 4))
