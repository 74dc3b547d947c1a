"""h-adaptive cubature controller and box rules (wanniernlqg_amd/cubature.py) on the CPU.

Pinned by: the reference's own known-answer cases of test/response_integration_unit.jl (values,
evaluation counts, budget and error behaviour, restated as Python calls); the rules' polynomial degrees
(Gauss-Kronrod 7/15 exact to degree 13 / 22, Genz-Malik exact to degree 7 with an embedded degree-5
rule, neither exact one degree higher); a sharply peaked Gaussian with a closed-form integral; and
invariants of the final partition (the boxes tile the domain and re-sum to the value)."""
import itertools
import math

import numpy as np
import pytest

from wanniernlqg_amd import cubature as C


def test_gauss_kronrod_degrees():
    r = C.gauss_kronrod()
    x = r.nodes[:, 0]
    for k in range(0, 30):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        kr = float(np.dot(r.weights, x ** k))
        ga = float(np.dot(r.embedded, x ** k))
        if k <= 22:
            assert abs(kr - exact) < 1e-14, k
        if k <= 13:
            assert abs(ga - exact) < 1e-14, k
        if k == 14:
            assert abs(ga - exact) > 1e-6
        if k == 24:
            assert abs(kr - exact) > 1e-9


@pytest.mark.parametrize("d", [2, 3])
def test_genz_malik_degrees(d):
    r = C.genz_malik(d)
    assert r.nodes.shape[0] == C.rule_nodes(d) == (17 if d == 2 else 33)
    rng = np.random.default_rng(d)
    lo = rng.uniform(-1, 0, d)
    hi = lo + rng.uniform(0.5, 2, d)
    half = 0.5 * (hi - lo)
    pts = 0.5 * (lo + hi) + half * r.nodes
    vol = np.prod(half)

    def exact(alpha):
        return np.prod([(hi[i] ** (a + 1) - lo[i] ** (a + 1)) / (a + 1) for i, a in enumerate(alpha)])

    for alpha in itertools.product(range(9), repeat=d):
        deg = sum(alpha)
        if deg > 8:
            continue
        f = np.prod(pts ** np.array(alpha), axis=1)
        ex = exact(alpha)
        hi7 = vol * np.dot(r.weights, f)
        lo5 = vol * np.dot(r.embedded, f)
        scale = max(1.0, abs(ex))
        if deg <= 7:
            assert abs(hi7 - ex) < 1e-12 * scale, alpha
        if deg <= 5:
            assert abs(lo5 - ex) < 1e-12 * scale, alpha
    # one degree higher breaks each rule: a pure eighth (sixth) power along one axis
    a8 = (8,) + (0,) * (d - 1)
    a6 = (6,) + (0,) * (d - 1)
    f8 = np.prod(pts ** np.array(a8), axis=1)
    f6 = np.prod(pts ** np.array(a6), axis=1)
    assert abs(vol * np.dot(r.weights, f8) - exact(a8)) > 1e-8
    assert abs(vol * np.dot(r.embedded, f6) - exact(a6)) > 1e-8


def test_reference_multifrequency_complex_case():
    dom = C.IntegrationDomain((-1.0, 2.0), (3.0, 5.0), "cartesian")
    alg = C.HAdaptiveCubature(atol=1e-11, rtol=1e-10, max_evaluations=1000)

    def ev(x):
        out = np.zeros((x.shape[0], 2, 2), dtype=np.complex128)
        out[:, 0, 0] = x[:, 0] ** 2 + 1j * x[:, 1]
        out[:, 0, 1] = 2 * x[:, 0] * x[:, 1]
        out[:, 1, 0] = 3
        out[:, 1, 1] = x[:, 1]
        return out

    expected = np.array([[28 + 42j, 84], [36, 42]])
    integ = C.integrate_response(ev, dom, alg, normalization="integral")
    assert np.max(np.abs(integ.value - expected)) < 1e-10
    assert integ.estimate_converged and integ.termination == "estimated_tolerance"
    assert integ.evaluations == 17
    assert integ.domain_measure == 12
    assert integ.normalization == "integral"
    mean = C.integrate_response(ev, dom, alg)
    assert np.max(np.abs(mean.value - expected / 12)) < 1e-10


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("divisions", [1, 2])
def test_reference_complete_rules(d, divisions):
    lower = (-1.0, 2.0, -2.0)[:d]
    upper = (3.0, 5.0, 4.0)[:d]
    dom = C.IntegrationDomain(lower, upper, "cartesian")
    mid = (np.array(lower) + np.array(upper)) / 2
    measure = float(np.prod(np.subtract(upper, lower)))

    def ev(x):
        out = np.zeros((x.shape[0], 2, 2, 2), dtype=np.complex128)
        out[:, 0, 0, 0] = x[:, 0] ** 2 + 1j * x.sum(axis=1)
        out[:, 0, 1, 0] = 2 * np.prod(x, axis=1)
        out[:, 1, 0, 0] = 3
        out[:, 1, 1, 0] = x[:, d - 1]
        out[..., 1] = (2 - 1j) * out[..., 0]
        return out

    exp = np.zeros((2, 2, 2), dtype=np.complex128)
    exp[0, 0, 0] = 7 / 3 + 1j * mid.sum()
    exp[0, 1, 0] = 2 * np.prod(mid)
    exp[1, 0, 0] = 3
    exp[1, 1, 0] = mid[d - 1]
    exp[..., 1] = (2 - 1j) * exp[..., 0]
    count = (15, 17, 33)[d - 1] * divisions ** d
    alg = C.HAdaptiveCubature(atol=1e-11, rtol=1e-10, max_evaluations=count, initial_divisions=divisions)
    mean = C.integrate_response(ev, dom, alg)
    assert np.max(np.abs(mean.value - exp)) < 1e-10
    assert mean.evaluations == count and mean.domain_measure == measure and mean.estimate_converged
    integ = C.integrate_response(ev, dom, alg, normalization="integral")
    assert np.max(np.abs(integ.value - measure * exp)) < 1e-9
    called = []
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: called.append(1) or np.ones(x.shape[0]), dom,
                             C.HAdaptiveCubature(max_evaluations=count - 1, initial_divisions=divisions))
    assert not called


def test_reference_budget_and_errors():
    dom = C.IntegrationDomain((-1.0, 2.0), (3.0, 5.0), "cartesian")
    res = C.integrate_response(lambda x: np.exp(3 * x[:, 0] + x[:, 1]), dom,
                               C.HAdaptiveCubature(atol=0, rtol=1e-14, max_evaluations=17))
    assert res.termination == "evaluation_budget" and not res.estimate_converged
    assert math.isfinite(float(res.value))
    with pytest.raises(ValueError):
        C.HAdaptiveCubature(atol=0, rtol=0)
    with pytest.raises(ValueError):
        C.HAdaptiveCubature(max_evaluations=0)
    with pytest.raises(ValueError):
        C.HAdaptiveCubature(initial_divisions=True)
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: np.ones(x.shape[0]), C.IntegrationDomain((0.0, 0.0), (1e200, 1e200)),
                             C.HAdaptiveCubature())
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: np.ones(x.shape[0]), C.IntegrationDomain((0.0, 0.0), (1e-200, 1e-200)),
                             C.HAdaptiveCubature())
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: np.ones(x.shape[0]), dom, C.HAdaptiveCubature(max_evaluations=16))
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: np.ones(x.shape[0]), dom,
                             C.HAdaptiveCubature(initial_divisions=2, max_evaluations=67))
    with pytest.raises(ValueError):
        C.integrate_response(lambda x: np.full(x.shape[0], np.nan), dom, C.HAdaptiveCubature())
    calls = [0]

    def growing(x):
        calls[0] += 1
        return np.repeat(np.exp(4 * x[:, :1]), calls[0], axis=1)

    with pytest.raises(ValueError):
        C.integrate_response(growing, dom, C.HAdaptiveCubature(atol=0, rtol=1e-15))
    assert calls[0] == 2
    with pytest.raises(ValueError):
        C.IntegrationDomain((0.0,), (1.0, 2.0))
    with pytest.raises(ValueError):
        C.IntegrationDomain((0.0, 0, 0, 0), (1.0, 1, 1, 1))
    with pytest.raises(ValueError):
        C.IntegrationDomain((1.0,), (0.5,))


def _peak(x0, s):
    def f(x):
        return np.exp(-np.sum((x - x0) ** 2, axis=1) / (2 * s * s))

    exact = np.prod([s * math.sqrt(math.pi / 2) * (math.erf((1 - c) / (math.sqrt(2) * s)) +
                                                   math.erf(c / (math.sqrt(2) * s))) for c in x0])
    return f, exact


def test_peaked_gaussian_closed_form():
    """A 3D Gaussian of width 0.04 inside the unit cube: the estimate is met and the true error lies
    below it; many boxes per step reach the same accuracy with about as many evaluations in far fewer
    evaluator calls."""
    f, exact = _peak(np.array([0.31, 0.62, 0.47]), 0.04)
    alg = C.HAdaptiveCubature(atol=0, rtol=1e-6, max_evaluations=1_000_000)
    one = C.integrate_response(f, C.IntegrationDomain(), alg, keep_boxes=True)
    many = C.integrate_response(f, C.IntegrationDomain(), alg, boxes_per_step=16, keep_boxes=True)
    for res in (one, many):
        assert res.estimate_converged
        assert abs(res.value - exact) < 1e-6 * exact
        assert abs(res.value - exact) < res.embedded_estimate
        # the final boxes tile the unit cube and re-sum to the value
        vols = np.prod(res.boxes_upper - res.boxes_lower, axis=1)
        assert abs(vols.sum() - 1.0) < 1e-12
        again = C.resum_boxes(f, res.boxes_lower, res.boxes_upper)
        assert abs(again - res.value) < 1e-14 * abs(res.value)
    assert many.evaluations < 1.05 * one.evaluations
    assert many.steps * 10 < one.steps


def test_split_axis_follows_the_fourth_difference():
    r = C.genz_malik(3)
    lo = np.zeros((1, 3))
    hi = np.ones((1, 3))
    for axis in range(3):
        nodes = C.box_nodes(r, lo, hi)[0]
        vals = (nodes[:, axis] ** 4 + 0.1 * nodes[:, (axis + 1) % 3] ** 2)[None, :]
        _, _, k = C.apply_rule(r, lo, hi, vals)
        assert k[0] == axis
    # a function without fourth-order variation: ties go to the widest axis
    hi2 = np.array([[1.0, 3.0, 2.0]])
    vals = np.ones((1, 33))
    _, _, k = C.apply_rule(r, lo, hi2, vals)
    assert k[0] == 1


def test_heap_pops_in_descending_order():
    rng = np.random.default_rng(0)
    h = C._MaxHeap()
    es = rng.integers(0, 40, 300).astype(float)      # many ties
    for e in es:
        h.push(C._Box(None, None, None, float(e), 0))
    out = [h.pop().E for _ in range(len(es))]
    assert out == sorted(es, reverse=True)
