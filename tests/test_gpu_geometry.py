"""GPU parity of the BCD and intrinsic AHC stages against the oracle (oracle/geometry.py), through
the C ABI.  Bar: FP64, 1e-10 relative (Frobenius norm over all components and Fermi energies).
Each test prints the case it runs before the GPU call."""
import math
import sys

import numpy as np
import pytest

from oracle import geometry as G
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import TightBindingModel, synthetic_model

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def qwz_layers(mass, a=2.5, c=3.0):
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    rv = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]
    hr = np.zeros((5, 2, 2), dtype=complex)
    hr[0] = mass * sz
    hr[1] = sx / 2j + sz / 2
    hr[2] = -sx / 2j + sz / 2
    hr[3] = sy / 2j + sz / 2
    hr[4] = -sy / 2j + sz / 2
    return TightBindingModel(np.diag([a, a, c]), np.array(rv), np.ones(5, dtype=np.int32), hr, None)


BCD_CASES = {
    "n6_thermal": (lambda: synthetic_model(6, 1, 8), dict(fermi_energies=[-0.5, 0.0, 0.7], temperature=600.0)),
    "n8_pos_gauss": (lambda: synthetic_model(8, 1, 5, position_scale=0.1),
                     dict(fermi_energies=[-0.3, 0.4], temperature=0.0, fermi_surface_broadening=0.2)),
    "n32_pos_eta": (lambda: synthetic_model(32, 1, 11, position_scale=0.1),
                    dict(fermi_energies=[-1.0, 0.0, 1.0], temperature=300.0, denominator_regularization=1e-3)),
}


def bcd_engine(model, kw, **extra):
    t = dict(berry_curvature_dipole=True, denominator_regularization=0.0, degeneracy_threshold=1e-9)
    t.update(kw)
    t.update(extra)
    return E.Engine(model, E.EngineTask(**t))


def bcd_task(kw):
    return G.GeometryTask(fermi_energies=kw["fermi_energies"], temperature=kw.get("temperature", 0.0),
                          fs_broadening=kw.get("fermi_surface_broadening") or None,
                          denominator_regularization=kw.get("denominator_regularization", 0.0),
                          degeneracy_threshold=1e-9)


@pytest.mark.parametrize("case", sorted(BCD_CASES))
def test_bcd_points(case):
    make, kw = BCD_CASES[case]
    m = make()
    k = np.random.default_rng(4).uniform(-0.5, 1.0, (9, 3))
    say(f"[bcd points] {case}")
    with bcd_engine(m, kw) as eng:
        got = eng.eval_points(k)["bcd"]
    ref = G.bcd_point_moments(m, bcd_task(kw), k)
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-10, rel(got, ref)


@pytest.mark.parametrize("case,backend", [("n6_thermal", "planes"), ("n8_pos_gauss", "direct"),
                                          ("n8_pos_gauss", "planes"), ("n32_pos_eta", "auto")])
def test_bcd_mesh(case, backend):
    make, kw = BCD_CASES[case]
    m = make()
    mesh = (5, 4, 6)
    say(f"[bcd mesh] {case} {backend} {mesh}")
    with bcd_engine(m, kw, fourier_backend=backend) as eng:
        lanes, total, st = eng.run_mesh(mesh)
    rl, rt = G.integrate_geometry_mesh(m, bcd_task(kw), mesh, "bcd")
    assert lanes["bcd"].shape == rl.shape
    assert rel(lanes["bcd"], rl) < 1e-10
    assert rel(total["bcd"], rt) < 1e-10


AHC_CASES = {
    "n6": (lambda: synthetic_model(6, 1, 3), [-0.5, 0.0, 0.5]),
    "n12_pos": (lambda: synthetic_model(12, 1, 9, position_scale=0.1), [-1.0, -0.2, 0.3, 1.5]),
    "n32_pos": (lambda: synthetic_model(32, 1, 2, position_scale=0.1), [0.0]),
}


@pytest.mark.parametrize("case", sorted(AHC_CASES))
def test_ahc_points_and_mesh(case):
    make, mus = AHC_CASES[case]
    m = make()
    k = np.random.default_rng(6).uniform(0, 1, (7, 3))
    say(f"[ahc] {case}")
    task = G.GeometryTask(fermi_energies=mus)
    with E.Engine(m, E.EngineTask(intrinsic_hall=True, fermi_energies=mus, denominator_regularization=0.0)) as eng:
        got = eng.eval_points(k)["ahc"]
        lanes, total, _ = eng.run_mesh((4, 5, 3))
    assert rel(got, G.ahc_point_values(m, task, k)) < 1e-10
    rl, rt = G.integrate_geometry_mesh(m, task, (4, 5, 3), "ahc")
    assert rel(lanes["ahc"], rl) < 1e-10
    assert rel(total["ahc"], rt) < 1e-10


@pytest.mark.parametrize("mass", [1.0, -1.2, 2.6])
def test_ahc_chern_layers_quantized(mass):
    """Stacked Qi-Wu-Zhang layers: sigma_xy = -(e^2/h) C / c (C = +-1 for 0 < |m| < 2, else 0)."""
    c = 3.0
    m = qwz_layers(mass, c=c)
    chern = 0 if abs(mass) > 2 else (-1 if mass > 0 else 1)
    say(f"[ahc chern] mass {mass}")
    with E.Engine(m, E.EngineTask(intrinsic_hall=True, fermi_energies=[0.0], denominator_regularization=0.0)) as eng:
        _, total, _ = eng.run_mesh((48, 48, 1))
    quantum = G.RESPONSE_CHARGE_C ** 2 / (2 * math.pi * G.RESPONSE_HBAR_JS) / (c * 1e-10)
    sxy = total["ahc"][0, 0, 1]
    assert abs(sxy - (-quantum * chern)) < 1e-6 * quantum, (sxy / quantum, chern)


def test_bcd_gap_error():
    base = synthetic_model(2, 1, 4)
    hr = np.zeros((base.num_r_vectors, 4, 4), dtype=complex)
    hr[:, :2, :2] = base.hamiltonian_r
    hr[:, 2:, 2:] = base.hamiltonian_r
    dup = TightBindingModel(base.lattice, base.r_vectors, base.r_degeneracies, hr, None)
    say("[bcd gap] degenerate copies")
    with E.Engine(dup, E.EngineTask(berry_curvature_dipole=True, fermi_energies=[0.0], temperature=300.0,
                                    denominator_regularization=0.0)) as eng:
        with pytest.raises(E.WnlqgError) as ei:
            eng.run_mesh((3, 3, 3))
    assert ei.value.code == -3


def test_bcd_and_ahc_in_one_engine_match_separate_engines():
    m = synthetic_model(10, 1, 4, position_scale=0.1)
    mus = [-0.4, 0.2]
    common = dict(fermi_energies=mus, temperature=0.0, fermi_surface_broadening=0.1, denominator_regularization=0.0,
                  degeneracy_threshold=1e-9)
    k = np.random.default_rng(8).uniform(0, 1, (6, 3))
    say("[bcd+ahc] one engine")
    with E.Engine(m, E.EngineTask(berry_curvature_dipole=True, intrinsic_hall=True, **common)) as both, \
            E.Engine(m, E.EngineTask(berry_curvature_dipole=True, **common)) as b, \
            E.Engine(m, E.EngineTask(intrinsic_hall=True, **common)) as a:
        pb = both.eval_points(k)
        lb, tb, _ = both.run_mesh((4, 3, 5))
        assert np.array_equal(pb["bcd"], b.eval_points(k)["bcd"])
        assert np.array_equal(pb["ahc"], a.eval_points(k)["ahc"])
        _, t1, _ = b.run_mesh((4, 3, 5))
        _, t2, _ = a.run_mesh((4, 3, 5))
    assert np.array_equal(tb["bcd"], t1["bcd"]) and np.array_equal(tb["ahc"], t2["ahc"])


def test_bcd_periodic_model_matches_reference_closed_form():
    """The reference's periodic two-band model (test/berry_multipole_periodic_unit.jl:1-68) on the engine:
    every moment component against the closed-form curvature and band gradients (no oracle in between)."""
    from test_oracle_response_pins import (MU_PERIODIC, T_PERIODIC, periodic_bcd_closed_form,
                                           periodic_bcd_model)
    say("bcd periodic two-band model, 32x32x1, mu -1.3 eV, T 800 K")
    model = periodic_bcd_model()
    with bcd_engine(model, dict(fermi_energies=[MU_PERIODIC], temperature=T_PERIODIC),
                    degeneracy_threshold=0.002) as eng:
        _, total, _ = eng.run_mesh((32, 32, 1))
    ref = periodic_bcd_closed_form(32)
    got = total["bcd"][0]
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_bcd_ahc_run_points_against_oracle():
    """Weighted point list on a BCD + AHC engine: each lane is the weighted sum of the oracle's per-point
    values over its items, the total that sum divided by the measure."""
    m = synthetic_model(8, 1, 5, position_scale=0.1)
    mus = [-0.3, 0.4]
    rng = np.random.default_rng(12)
    k = rng.uniform(-0.5, 1.0, (301, 3))
    w = rng.uniform(0.1, 2.0, 301)
    nl = 7
    say(f"[geometry run_points] n=8, m=301, lanes={nl}")
    with E.Engine(m, E.EngineTask(berry_curvature_dipole=True, intrinsic_hall=True, fermi_energies=mus,
                                  temperature=0.0, fermi_surface_broadening=0.2, denominator_regularization=0.0,
                                  degeneracy_threshold=1e-9)) as eng:
        lanes, total = eng.run_points(k, w, nl)
    task = G.GeometryTask(fermi_energies=mus, temperature=0.0, fs_broadening=0.2, degeneracy_threshold=1e-9)
    vals = {"bcd": G.bcd_point_moments(m, task, k), "ahc": G.ahc_point_values(m, task, k)}
    vol = abs(np.linalg.det(m.lattice))
    measure = {"bcd": vol, "ahc": vol * 1e-30}
    for key in ("bcd", "ahc"):
        for lane in range(nl):
            a, b = lane * 301 // nl, (lane + 1) * 301 // nl
            assert rel(lanes[key][lane], np.einsum("k,k...->...", w[a:b], vals[key][a:b])) < 1e-10, (key, lane)
        assert rel(total[key], np.einsum("k,k...->...", w, vals[key]) / measure[key]) < 1e-10, key


def test_bcd_run_points_over_a_shuffled_mesh_equals_run_mesh():
    """Every point of a 5 x 4 x 6 mesh, in a random order, with weights 1/N_k: the mesh total."""
    make, kw = BCD_CASES["n8_pos_gauss"]
    m = make()
    mesh = (5, 4, 6)
    nk = int(np.prod(mesh))
    idx = np.arange(nk)
    pts = np.stack([idx % 5 / 5.0, (idx // 5) % 4 / 4.0, idx // 20 / 6.0], axis=1)
    perm = np.random.default_rng(3).permutation(nk)
    say("[bcd run_points] shuffled 5x4x6 mesh")
    with bcd_engine(m, kw) as eng:
        _, mesh_total, _ = eng.run_mesh(mesh)
        _, pts_total = eng.run_points(pts[perm], np.full(nk, 1.0 / nk), 9)
    assert rel(pts_total["bcd"], mesh_total["bcd"]) < 1e-12
