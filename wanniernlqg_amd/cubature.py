"""h-adaptive cubature over a box of the Brillouin zone, with every rule's nodes evaluated by the GPU
engine in one batch.

Mirrors the reference's adaptive response integration (``integrate_response`` with
``HAdaptiveCubature``, ``src/Runtime/Execution/ResponseIntegration.jl:100-155``; the tree controller
``_response_cubature_tree``, ``src/Runtime/Execution/ResponseCubatureTree.jl:20-106``; the algorithm
and domain types, ``src/Runtime/IntegrationAlgorithms.jl``).  The reference takes its box rules from
the HCubature.jl 1.8.0 package, which is not part of the reference tree; the rules are restated here
from their published definitions:

* 1D: the 7-point Gauss / 15-point Kronrod pair (Piessens et al., QUADPACK ``qk15``): value = Kronrod
  sum, estimate = |Kronrod - Gauss|, split at the midpoint;
* 2D / 3D: the Genz-Malik degree-7 rule with its embedded degree-5 rule (Genz & Malik, J. Comput. Appl.
  Math. 6 (1980) 295-302): 1 + 4d + 2d(d-1) + 2^d nodes (17 in 2D, 33 in 3D), estimate = norm of the
  difference, split axis = the axis with the largest fourth divided difference
  |f(c+l3 e_i) + f(c-l3 e_i) + 12 f(c) - 7 (f(c+l2 e_i) + f(c-l2 e_i))|; differences equal to a relative
  1e-10 (rounding noise) go to the widest axis, then the lowest index.

The tree is the reference's controller: a binary max-heap of boxes keyed on their estimates, the
initial ``initial_divisions^d`` partition, pop the worst box, halve it along its split axis, update the
running value and estimate, stop when estimate <= max(rtol |value|, atol) (|.| the maximum absolute
complex entry of the tensor) or the evaluation budget is spent, then re-sum over the heap.  One
MI355X-specific extension: ``boxes_per_step`` pops up to that many boxes per step and evaluates all
their children's nodes in one engine call (the reference refines one box per step; with the default
``boxes_per_step = 1`` the sequence of boxes and the arithmetic are the reference's).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

# ---------------------------------------------------------------------------------------------
# algorithm and domain types (IntegrationAlgorithms.jl)


@dataclass(frozen=True)
class HAdaptiveCubature:
    """Controls of the h-adaptive backend (``HAdaptiveCubature``, ``IntegrationAlgorithms.jl:50-80``)."""
    atol: float = 1e-8
    rtol: float = 1e-5
    max_evaluations: int = 100_000
    initial_divisions: int = 1

    def __post_init__(self):
        for v in (self.atol, self.rtol):
            if isinstance(v, bool) or not isinstance(v, (int, float, np.integer, np.floating)):
                raise ValueError("finite nonnegative atol/rtol with at least one positive value required")
        atol, rtol = float(self.atol), float(self.rtol)
        if not (math.isfinite(atol) and math.isfinite(rtol) and atol >= 0 and rtol >= 0 and (atol > 0 or rtol > 0)):
            raise ValueError("finite nonnegative atol/rtol with at least one positive value required")
        for name in ("initial_divisions", "max_evaluations"):
            v = getattr(self, name)
            if isinstance(v, bool) or not isinstance(v, (int, np.integer)) or v <= 0:
                raise ValueError(f"{name} must be positive")
        object.__setattr__(self, "atol", atol)
        object.__setattr__(self, "rtol", rtol)


@dataclass(frozen=True)
class IntegrationDomain:
    """Finite 1- to 3-dimensional box in fractional or Cartesian (1/Angstrom) coordinates
    (``IntegrationDomain``, ``IntegrationAlgorithms.jl:62-97``).  The default is the full zone [0, 1]^3."""
    lower: tuple = (0.0, 0.0, 0.0)
    upper: tuple = (1.0, 1.0, 1.0)
    coordinates: str = "fractional"

    def __post_init__(self):
        lo, hi = tuple(float(x) for x in self.lower), tuple(float(x) for x in self.upper)
        if len(lo) != len(hi):
            raise ValueError("matching integration bounds required")
        if not 1 <= len(lo) <= 3:
            raise ValueError("integration dimension must be one, two or three")
        if not (all(math.isfinite(x) for x in lo + hi) and all(a < b for a, b in zip(lo, hi))):
            raise ValueError("finite ordered integration domain required")
        if self.coordinates not in ("fractional", "cartesian"):
            raise ValueError("coordinates must be fractional or cartesian")
        object.__setattr__(self, "lower", lo)
        object.__setattr__(self, "upper", hi)

    @property
    def dimension(self) -> int:
        return len(self.lower)

    @property
    def measure(self) -> float:
        return math.prod(b - a for a, b in zip(self.lower, self.upper))


def rule_nodes(dimension: int) -> int:
    """Nodes of one complete box rule (``integration_rule_nodes``, ``IntegrationAlgorithms.jl:99-100``)."""
    d = int(dimension)
    return 15 if d == 1 else 1 + 4 * d + 2 * d * (d - 1) + 2 ** d


def validate_budget(domain: IntegrationDomain, algorithm: HAdaptiveCubature) -> int:
    """The initial partition's complete rules must fit the budget before anything is evaluated."""
    need = rule_nodes(domain.dimension) * algorithm.initial_divisions ** domain.dimension
    if need > algorithm.max_evaluations:
        raise ValueError(f"max_evaluations={algorithm.max_evaluations} is below the initial partition's "
                         f"{need} rule nodes")
    return need


# ---------------------------------------------------------------------------------------------
# box rules


# QUADPACK qk15: Kronrod abscissae (descending, the last is 0), Kronrod weights, Gauss weights of the
# 7-point rule at the odd-indexed abscissae
_XGK = (0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
        0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
        0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
        0.207784955007898467600689403773245, 0.000000000000000000000000000000000)
_WGK = (0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
        0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
        0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
        0.204432940075298892414161999234649, 0.209482141084727828012999174891714)
_WG = (0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
       0.381830050505118944950369775488975, 0.417959183673469387755102040816327)


@dataclass
class BoxRule:
    """Unit nodes on [-1, 1]^d and the weight vectors of the rule and its embedded lower-order rule
    (weights already include the 2^d unit-cube volume; a box of half-widths h gets prod(h) times them)."""
    dimension: int
    nodes: np.ndarray        # [m, d]
    weights: np.ndarray      # [m]
    embedded: np.ndarray     # [m]
    l2: float = 0.0
    l3: float = 0.0


def genz_malik(d: int) -> BoxRule:
    l2, l3, l4, l5 = math.sqrt(9 / 70), math.sqrt(9 / 10), math.sqrt(9 / 10), math.sqrt(9 / 19)
    two = 2.0 ** d
    w = (two * (12824 - 9120 * d + 400 * d * d) / 19683, two * 980 / 6561, two * (1820 - 400 * d) / 19683,
         two * 200 / 19683, 6859 / 19683)
    we = (two * (729 - 950 * d + 50 * d * d) / 729, two * 245 / 486, two * (265 - 100 * d) / 1458,
          two * 25 / 729, 0.0)
    nodes: List[np.ndarray] = [np.zeros(d)]
    cls = [0]
    for lam, c in ((l2, 1), (l3, 2)):
        for i in range(d):
            for s in (1.0, -1.0):
                p = np.zeros(d)
                p[i] = s * lam
                nodes.append(p)
                cls.append(c)
    for i in range(d):
        for j in range(i + 1, d):
            for si in (1.0, -1.0):
                for sj in (1.0, -1.0):
                    p = np.zeros(d)
                    p[i], p[j] = si * l4, sj * l4
                    nodes.append(p)
                    cls.append(3)
    for bits in range(2 ** d):
        nodes.append(np.array([(-l5 if (bits >> i) & 1 else l5) for i in range(d)]))
        cls.append(4)
    cls = np.array(cls)
    return BoxRule(d, np.array(nodes), np.array([w[c] for c in cls]), np.array([we[c] for c in cls]), l2, l3)


def gauss_kronrod() -> BoxRule:
    x = np.array([-v for v in _XGK[:-1]] + [0.0] + list(reversed(_XGK[:-1])))
    wk = np.array(list(_WGK[:-1]) + [_WGK[-1]] + list(reversed(_WGK[:-1])))
    wg = np.zeros(15)
    for i in range(7):          # Gauss abscissae are the odd-indexed Kronrod ones
        if i % 2 == 1:
            wg[i] = wg[14 - i] = _WG[i // 2]
    wg[7] = _WG[3]
    return BoxRule(1, x[:, None], wk, wg)


def make_rule(d: int) -> BoxRule:
    if d == 1:
        return gauss_kronrod()
    if d in (2, 3):
        return genz_malik(d)
    raise ValueError("integration dimension must be one, two or three")


def _norm(v: np.ndarray) -> float:
    """Maximum absolute complex entry (``_integration_norm``, ``ResponseIntegration.jl:57-59``)."""
    return float(np.max(np.abs(v))) if np.ndim(v) else float(abs(v))


def apply_rule(rule: BoxRule, lower: np.ndarray, upper: np.ndarray, values: np.ndarray):
    """Value, estimate and split axis of boxes [B, d] from their node values [B, m, ...]."""
    half = 0.5 * np.abs(upper - lower)                                  # [B, d]
    vol = np.prod(half, axis=1)                                         # [B]
    tail = values.shape[2:]
    flat = values.reshape(values.shape[0], values.shape[1], -1)
    hi = np.einsum("m,bmt->bt", rule.weights, flat)
    lo = np.einsum("m,bmt->bt", rule.embedded, flat)
    I = (vol[:, None] * hi).reshape((values.shape[0],) + tail)
    Ie = (vol[:, None] * lo).reshape((values.shape[0],) + tail)
    B = values.shape[0]
    E = np.max(np.abs((I - Ie).reshape(B, -1)), axis=1) if I[0].size else np.zeros(B)
    if rule.dimension == 1:
        kdiv = np.zeros(B, dtype=np.int64)
    else:
        d = rule.dimension
        f1 = flat[:, 0]
        diffs = np.empty((B, d))
        for i in range(d):
            f2 = flat[:, 1 + 2 * i] + flat[:, 2 + 2 * i]
            f3 = flat[:, 1 + 2 * d + 2 * i] + flat[:, 2 + 2 * d + 2 * i]
            diffs[:, i] = np.max(np.abs(f3 + 12.0 * f1 - 7.0 * f2), axis=1)
        # largest fourth difference; equal to a relative 1e-10: the widest axis, then the lowest index
        near = diffs >= diffs.max(axis=1, keepdims=True) * (1.0 - 1e-10)
        kdiv = np.argmax(np.where(near, half, -np.inf), axis=1).astype(np.int64)
    return I, E, kdiv


def box_nodes(rule: BoxRule, lower: np.ndarray, upper: np.ndarray) -> np.ndarray:
    """Nodes c + h p of boxes [B, d]: [B, m, d]."""
    c = 0.5 * (lower + upper)
    h = 0.5 * np.abs(upper - lower)
    return c[:, None, :] + h[:, None, :] * rule.nodes[None, :, :]


# ---------------------------------------------------------------------------------------------
# the tree controller (ResponseCubatureTree.jl:20-106)


@dataclass
class _Box:
    a: np.ndarray
    b: np.ndarray
    I: np.ndarray
    E: float
    kdiv: int


class _MaxHeap:
    """Array binary max-heap on the box estimate: push appends and sifts up past strictly smaller
    parents; pop moves the last entry to the root and sifts it down past the strictly larger child
    (the right child only when strictly larger than the left)."""

    def __init__(self):
        self.tree: List[_Box] = []

    def __len__(self):
        return len(self.tree)

    def push(self, x: _Box):
        t = self.tree
        t.append(x)
        i = len(t) - 1
        while i > 0:
            p = (i - 1) >> 1
            if x.E > t[p].E:
                t[i] = t[p]
                i = p
            else:
                break
        t[i] = x

    def pop(self) -> _Box:
        t = self.tree
        top = t[0]
        last = t.pop()
        if t:
            n = len(t)
            i = 0
            while True:
                lc = 2 * i + 1
                if lc >= n:
                    break
                c = lc
                if lc + 1 < n and t[lc + 1].E > t[lc].E:
                    c = lc + 1
                if t[c].E > last.E:
                    t[i] = t[c]
                    i = c
                else:
                    break
            t[i] = last
        return top


@dataclass
class IntegralResult:
    """``ResponseIntegralResult`` (``ResponseIntegration.jl:142-154``) plus the final boxes."""
    value: np.ndarray
    embedded_estimate: float
    evaluations: int
    termination: str                 # "estimated_tolerance" | "evaluation_budget"
    estimate_converged: bool
    domain_measure: float
    normalization: str
    norm: str = "maximum_absolute_complex_entry"
    max_evaluations: int = 0
    steps: int = 0                   # evaluator calls
    boxes_lower: Optional[np.ndarray] = field(default=None, repr=False)
    boxes_upper: Optional[np.ndarray] = field(default=None, repr=False)


def integrate_response(evaluator: Callable[[np.ndarray], np.ndarray], domain: IntegrationDomain,
                       algorithm: HAdaptiveCubature, normalization: str = "domain_average",
                       boxes_per_step: int = 1, keep_boxes: bool = False) -> IntegralResult:
    """Integrate a batched evaluator (nodes [m, d] -> values [m, ...], any fixed tensor shape) over the
    domain (``integrate_response``, ``ResponseIntegration.jl:100-155``).  Non-finite values and changing
    shapes raise; exhausting the budget keeps the value and reports termination "evaluation_budget"."""
    if normalization not in ("integral", "domain_average"):
        raise ValueError("unknown integration normalization")
    if isinstance(boxes_per_step, bool) or int(boxes_per_step) < 1:
        raise ValueError("boxes_per_step must be a positive integer")
    measure = domain.measure
    if not (math.isfinite(measure) and measure > 0):
        raise ValueError("integration measure must be positive and representable")
    scale = 1.0 / measure if normalization == "domain_average" else 1.0
    if not (math.isfinite(scale) and scale > 0):
        raise ValueError("integration normalization must be representable in Float64")
    validate_budget(domain, algorithm)
    d = domain.dimension
    rule = make_rule(d)
    m = rule.nodes.shape[0]
    shape = [None]
    calls = [0]

    def rules(lower: np.ndarray, upper: np.ndarray):
        nodes = box_nodes(rule, lower, upper)
        vals = np.asarray(evaluator(nodes.reshape(-1, d)))
        calls[0] += 1
        if vals.shape[0] != nodes.shape[0] * m:
            raise ValueError("evaluator returned the wrong number of values")
        if vals.size == 0 or not np.all(np.isfinite(vals)):
            raise ValueError("nonfinite or empty integration value")
        if shape[0] is None:
            shape[0] = vals.shape[1:]
        elif vals.shape[1:] != shape[0]:
            raise ValueError("integrand shape changed")
        vals = scale * vals.reshape((nodes.shape[0], m) + vals.shape[1:])
        return apply_rule(rule, lower, upper, vals)

    atol, rtol = algorithm.atol, algorithm.rtol
    maxevals, initdiv = algorithm.max_evaluations, algorithm.initial_divisions
    a = np.array(domain.lower)
    b = np.array(domain.upper)
    delta = (b - a) / initdiv
    # the initial partition, column-major over the division index (first axis fastest)
    lows, highs = [], []
    for flat in range(initdiv ** d):
        c = [(flat // initdiv ** i) % initdiv for i in range(d)]
        lows.append([a[i] + c[i] * delta[i] for i in range(d)])
        highs.append([b[i] if c[i] == initdiv - 1 else a[i] + (c[i] + 1) * delta[i] for i in range(d)])
    if initdiv == 1:
        highs[0] = list(b)
    lows, highs = np.array(lows), np.array(highs)
    Is, Es, ks = rules(lows, highs)
    heap = _MaxHeap()
    I = Is[0].copy()
    E = float(Es[0])
    numevals = m
    heap.push(_Box(lows[0], highs[0], Is[0], float(Es[0]), int(ks[0])))
    for j in range(1, lows.shape[0]):
        I = I + Is[j]
        E = E + float(Es[j])
        numevals += m
        heap.push(_Box(lows[j], highs[j], Is[j], float(Es[j]), int(ks[j])))
    final_sum = True
    if E <= max(rtol * _norm(I), atol) or numevals >= maxevals:
        final_sum = False            # the reference returns the running sums here
    else:
        while True:
            room = max(1, -(-(maxevals - numevals) // (2 * m)))
            take = min(int(boxes_per_step), len(heap), room)
            popped = [heap.pop() for _ in range(take)]
            lo1, hi1, lo2, hi2 = [], [], [], []
            for box in popped:
                k = box.kdiv
                w = (box.b[k] - box.a[k]) * 0.5
                a2 = box.a.copy()
                a2[k] += w
                b2 = box.b.copy()
                b2[k] -= w
                lo1.append(a2); hi1.append(box.b)
                lo2.append(box.a); hi2.append(b2)
            lows = np.array([x for pair in zip(lo1, lo2) for x in pair])
            highs = np.array([x for pair in zip(hi1, hi2) for x in pair])
            Is, Es, ks = rules(lows, highs)
            stop = False
            for j, box in enumerate(popped):
                b1 = _Box(lows[2 * j], highs[2 * j], Is[2 * j], float(Es[2 * j]), int(ks[2 * j]))
                b2_ = _Box(lows[2 * j + 1], highs[2 * j + 1], Is[2 * j + 1], float(Es[2 * j + 1]),
                           int(ks[2 * j + 1]))
                heap.push(b1)
                heap.push(b2_)
                I = I + (b1.I + b2_.I - box.I)
                E = E + (b1.E + b2_.E - box.E)
                numevals += 2 * m
            inorm = _norm(I)
            if E <= max(rtol * inorm, atol) or numevals >= maxevals or not math.isfinite(inorm):
                stop = True
            if stop:
                break
    if final_sum:
        I = np.zeros_like(I)
        E = 0.0
        for box in heap.tree:
            I = I + box.I
            E = E + box.E
    if not (math.isfinite(E) and math.isfinite(_norm(I))):
        raise ValueError("nonfinite cubature result")
    converged = E <= max(atol, rtol * _norm(I))
    res = IntegralResult(value=I, embedded_estimate=float(E), evaluations=numevals,
                         termination="estimated_tolerance" if converged else "evaluation_budget",
                         estimate_converged=converged, domain_measure=measure, normalization=normalization,
                         max_evaluations=maxevals, steps=calls[0])
    if keep_boxes:
        res.boxes_lower = np.array([bx.a for bx in heap.tree])
        res.boxes_upper = np.array([bx.b for bx in heap.tree])
    return res


# ---------------------------------------------------------------------------------------------
# engine evaluator (ResponseNodeEvaluator, ResponsePointEvaluation.jl:48-64)


# physical constants of the drivers' measures (Core/Constants.jl, SecondHarmonicDriver.jl:136-139)
_CHARGE_C = 1.602176634e-19
_EPSILON0 = 8.8541878128e-12
_HBAR_JS = 1.054571817e-34


class EngineEvaluator:
    """Integration nodes -> k points (inactive axes zero; Cartesian nodes mapped to fractional k by
    k = L k_cart / (2 pi), ``reciprocal_cartesian_to_fractional``, ``Core/Coordinates.jl:15-20``) -> the
    engine's per-point values of all its states, stacked on a leading state axis: [m, n_states, ...].

    Each state is scaled by the measure its dense-mesh total carries, so that the domain average is
    directly comparable with ``Engine.run_mesh`` totals: SC / IC values already hold their prefactors;
    BCD moments are divided by |det L| and AHC by |det L| 1e-30 (``integrate_geometry`` measure of the
    geometry drivers); SHG chi and sigma take the factors e / (eps0 V) 1e12 and e^2 / (hbar V)
    (``SecondHarmonicDriver.jl:135-139``, the same factors its adaptive route applies per node)."""

    def __init__(self, engine, domain: IntegrationDomain):
        self.engine = engine
        self.domain = domain
        self.states = list(engine.states)
        self.points = 0
        vol = abs(float(np.linalg.det(np.asarray(engine.model.lattice, dtype=np.float64))))
        self.factors = []
        for s in self.states:
            if s in ("sc", "ic"):
                self.factors.append(1.0)
            elif s == "bcd":
                self.factors.append(1.0 / vol)
            elif s == "ahc":
                self.factors.append(1.0 / (vol * 1e-30))
            elif s == "shg":
                self.factors.append(np.array([_CHARGE_C / (_EPSILON0 * vol) * 1e12, _CHARGE_C ** 2 / (_HBAR_JS * vol)]))
            else:
                raise ValueError(f"no integration measure for state {s!r}")

    def kpoints(self, nodes: np.ndarray) -> np.ndarray:
        nodes = np.asarray(nodes, dtype=np.float64).reshape(-1, self.domain.dimension)
        k = np.zeros((nodes.shape[0], 3))
        k[:, :nodes.shape[1]] = nodes
        if self.domain.coordinates == "cartesian":
            k = k @ np.asarray(self.engine.model.lattice, dtype=np.float64).T / (2.0 * math.pi)
        return k

    def __call__(self, nodes: np.ndarray) -> np.ndarray:
        k = self.kpoints(nodes)
        self.points += k.shape[0]
        vals = self.engine.eval_points(k)
        return np.stack([vals[s] * f for s, f in zip(self.states, self.factors)], axis=1)


def integrate_engine(engine, algorithm: HAdaptiveCubature, domain: Optional[IntegrationDomain] = None,
                     boxes_per_step: int = 1, keep_boxes: bool = False) -> IntegralResult:
    """Domain-average adaptive integral of the engine's states over the full zone (or a box of it); the
    value has the layout of a dense-mesh total with a leading state axis."""
    domain = domain or IntegrationDomain()
    return integrate_response(EngineEvaluator(engine, domain), domain, algorithm, "domain_average",
                              boxes_per_step=boxes_per_step, keep_boxes=keep_boxes)


def resum_boxes(evaluator: Callable[[np.ndarray], np.ndarray], lower: np.ndarray, upper: np.ndarray,
                scale: float = 1.0) -> np.ndarray:
    """Sum of the box rules over given boxes, in the given order, with values from another evaluator
    (checks an adaptive result's final partition against independent point values)."""
    d = lower.shape[1]
    rule = make_rule(d)
    nodes = box_nodes(rule, lower, upper)
    vals = np.asarray(evaluator(nodes.reshape(-1, d)))
    vals = scale * vals.reshape((lower.shape[0], rule.nodes.shape[0]) + vals.shape[1:])
    I, _, _ = apply_rule(rule, lower, upper, vals)
    out = np.zeros_like(I[0])
    for j in range(I.shape[0]):
        out = out + I[j]
    return out
