"""Adaptive cubature over the Brillouin zone with the GPU engine evaluating every rule's nodes.

Checks: (1) the adaptive value equals the box rules of its final partition re-evaluated with the
oracle's point values (oracle/kloop.py, oracle/geometry.py), 1e-10 relative — the GPU values, node
mapping and rule arithmetic all enter; (2) many boxes per step agree with one box per step within the
two embedded estimates; (3) the adaptive value agrees with the engine's dense-mesh integral of the same
smooth integrand (a different code path) to the adaptive tolerance."""
import os
import sys

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kloop
from wanniernlqg_amd import cubature as C
from wanniernlqg_amd import engine as E
from wanniernlqg_amd.model import read_wannier_tb, synthetic_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def maxrel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_adaptive_shift_injection_current():
    model = read_wannier_tb(FIXTURE)
    omega = np.linspace(0.0, 4.0, 40)
    kw = dict(photon_energies=omega, broadening=0.03, fermi_energy=0.1, shift_current=True, injection_current=True)
    alg = C.HAdaptiveCubature(atol=0.0, rtol=1e-4, max_evaluations=40_000)
    say("[adaptive sc/ic] fixture model, rtol 1e-4, one box per step")
    with E.Engine(model, E.EngineTask(**kw)) as eng:
        one = C.integrate_engine(eng, alg, keep_boxes=True)
        say(f"  evaluations {one.evaluations} steps {one.steps} estimate {one.embedded_estimate:.3e}")
        many = C.integrate_engine(eng, alg, boxes_per_step=32)
        say(f"  32 boxes per step: evaluations {many.evaluations} steps {many.steps}")
        _, dense, _ = eng.run_mesh((48, 48, 48))
    assert one.value.shape == (2, 40, 3, 3, 3)
    task = kloop.Task(photon_energies=omega, broadening=0.03, fermi_energy=0.1, shift_current=True,
                      injection_current=True)

    def oracle_eval(nodes):
        v = kloop.point_values(model, task, nodes)
        return np.stack([v["sc"], v["ic"]], axis=1)

    again = C.resum_boxes(oracle_eval, one.boxes_lower, one.boxes_upper)
    r = rel(one.value, again)
    say(f"  final partition re-evaluated by the oracle: rel {r:.3e} ({one.boxes_lower.shape[0]} boxes)")
    assert r < 1e-10
    assert one.estimate_converged and many.estimate_converged
    assert np.max(np.abs(one.value - many.value)) <= one.embedded_estimate + many.embedded_estimate
    for i, key in enumerate(("sc", "ic")):
        d = maxrel(one.value[i], dense[key])
        say(f"  {key}: adaptive vs dense 48^3 mesh max-entry rel {d:.3e}")
        assert d < 1e-3


def test_adaptive_berry_curvature_dipole_box():
    """BCD moments over a box of the zone in Cartesian coordinates (k = L k_cart / 2 pi)."""
    model = synthetic_model(6, 1, 8)
    kw = dict(fermi_energies=[-0.5, 0.0, 0.7], temperature=600.0)
    lat = np.asarray(model.lattice)
    b = 2 * np.pi * np.linalg.inv(lat).T          # rows: reciprocal vectors
    lo = 0.1 * np.diag(b)
    hi = 0.6 * np.diag(b)
    dom = C.IntegrationDomain(tuple(lo), tuple(hi), "cartesian")
    alg = C.HAdaptiveCubature(atol=0.0, rtol=5e-3, max_evaluations=8_000)
    say("[adaptive bcd] n6 thermal, Cartesian box, 8 boxes per step")
    t = dict(berry_curvature_dipole=True, denominator_regularization=0.0, degeneracy_threshold=1e-9, **kw)
    with E.Engine(model, E.EngineTask(**t)) as eng:
        ev = C.EngineEvaluator(eng, dom)
        res = C.integrate_response(ev, dom, alg, boxes_per_step=8, keep_boxes=True)
    say(f"  evaluations {res.evaluations} steps {res.steps} points {ev.points}")
    assert ev.points == res.evaluations
    task = G.GeometryTask(fermi_energies=kw["fermi_energies"], temperature=600.0, fs_broadening=None,
                          denominator_regularization=0.0, degeneracy_threshold=1e-9)

    vol = abs(np.linalg.det(lat))

    def oracle_eval(nodes):
        # the oracle's dense-mesh BCD total is the mean moment divided by |det L|
        k = nodes @ lat.T / (2 * np.pi)
        return G.bcd_point_moments(model, task, k)[:, None] / vol

    again = C.resum_boxes(oracle_eval, res.boxes_lower, res.boxes_upper, scale=1.0 / dom.measure)
    r = rel(res.value, again)
    say(f"  final partition re-evaluated by the oracle: rel {r:.3e}")
    assert r < 1e-10
