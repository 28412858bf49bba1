"""Truncation policies: ``coeff``, ``weight``, ``topn`` and ``approx_topn``, composed with ``&`` and ``|``.

Mirrors the reference's factories (crates/paulistrings-py/src/truncation.rs:11-33) and their semantics
(crates/paulistrings/src/truncation/builtin.rs):

* ``coeff(eps)`` -- CoefficientThreshold: keep iff ``eps < 0`` or ``|c|^2 > eps^2`` (builtin.rs:19-41);
* ``weight(k)`` -- WeightCutoff: keep iff the Pauli weight is at most ``k`` (builtin.rs:43-59);
* ``topn(n)`` -- TopN: after every layer keep the ``n`` largest ``|c|``, ties at the boundary in canonical
  order (builtin.rs:61-150);
* ``approx_topn(n)`` -- ApproxTopN: after every layer keep ``n`` terms by octave of ``|c|^2``, whole octaves from
  the top and the boundary octave in canonical order (builtin.rs:241-468);
* ``a & b`` -- And: both filters, both layer passes (builtin.rs:470-526);
* ``a | b`` -- Or: either filter; an operand with a layer pass is rejected (builtin.rs:568-626).

A policy lowers to the C ABI's ``pp_policy``: a postfix program of at most 15 filter nodes and one layer pass
(TopN or ApproxTopN).
"""

from __future__ import annotations

from dataclasses import dataclass

from . import _native

NODE_TRUE, NODE_COEFF, NODE_WEIGHT, NODE_AND, NODE_OR = 0, 1, 2, 3, 4


class Truncation:
    """A built-in truncation policy tree."""

    def __and__(self, other: "Truncation") -> "Truncation":
        return And(self, _check(other))

    def __or__(self, other: "Truncation") -> "Truncation":
        other = _check(other)
        for side, name in ((self, "its first operand has"), (other, "its second operand has")):
            if side.has_pass():
                raise ValueError("Or combines per-term filters only, but " + name +
                                 " a layer pass (TopN or ApproxTopN); use And to add a filter to a layer pass")
        return Or(self, other)

    # -- lowering --
    def nodes(self) -> list[tuple[int, int, float]]:
        raise NotImplementedError

    def topn_values(self) -> list[int]:
        return []

    def approx_topn_values(self) -> list[int]:
        return []

    def has_pass(self) -> bool:
        return bool(self.topn_values()) or bool(self.approx_topn_values())

    def lower(self) -> "_native.Policy":
        nodes = self.nodes()
        if len(nodes) > 15:
            raise ValueError(f"the per-term filter has {len(nodes)} nodes; at most 15 lower to the device")
        tops, approx = self.topn_values(), self.approx_topn_values()
        if tops and approx:
            raise ValueError("a policy with both a TopN and an ApproxTopN layer pass does not lower to the device")
        # TopN(a) then TopN(b) keeps the min(a, b) first by magnitude rank, ties taken in the same order; the same
        # holds for ApproxTopN by octave rank (a prefix of a stable order is that order's prefix).
        passes = tops or approx
        p = _native.Policy()
        p.n_nodes = len(nodes)
        for i, (op, k, eps) in enumerate(nodes):
            p.nodes[i].op = op
            p.nodes[i].k = k
            p.nodes[i].eps = eps
        p.has_topn = 0 if not passes else (1 if tops else 2)
        p.topn = int(min(passes)) if passes else 0
        return p


def _check(p) -> Truncation:
    if not isinstance(p, Truncation):
        raise TypeError(f"expected a truncation policy, got {type(p).__name__}")
    return p


@dataclass(frozen=True)
class Coeff(Truncation):
    eps: float

    def nodes(self):
        return [(NODE_COEFF, 0, float(self.eps))]

    def __repr__(self):
        return f"coeff({self.eps})"


@dataclass(frozen=True)
class Weight(Truncation):
    k: int

    def nodes(self):
        return [(NODE_WEIGHT, int(self.k), 0.0)]

    def __repr__(self):
        return f"weight({self.k})"


@dataclass(frozen=True)
class TopN(Truncation):
    n: int

    def nodes(self):
        return []

    def topn_values(self):
        return [int(self.n)]

    def __repr__(self):
        return f"topn({self.n})"


@dataclass(frozen=True)
class ApproxTopN(Truncation):
    n: int

    def nodes(self):
        return []

    def approx_topn_values(self):
        return [int(self.n)]

    def __repr__(self):
        return f"approx_topn({self.n})"


@dataclass(frozen=True)
class And(Truncation):
    a: Truncation
    b: Truncation

    def nodes(self):
        na, nb = self.a.nodes(), self.b.nodes()
        if not na:
            return nb
        if not nb:
            return na
        return na + nb + [(NODE_AND, 0, 0.0)]

    def topn_values(self):
        return self.a.topn_values() + self.b.topn_values()

    def approx_topn_values(self):
        return self.a.approx_topn_values() + self.b.approx_topn_values()

    def __repr__(self):
        return f"({self.a!r} & {self.b!r})"


@dataclass(frozen=True)
class Or(Truncation):
    a: Truncation
    b: Truncation

    def nodes(self):
        na, nb = self.a.nodes(), self.b.nodes()
        # an operand with no filter keeps everything: so does the Or
        if not na or not nb:
            return [(NODE_TRUE, 0, 0.0)]
        return na + nb + [(NODE_OR, 0, 0.0)]

    def __repr__(self):
        return f"({self.a!r} | {self.b!r})"


def coeff(eps: float) -> Truncation:
    """Drop terms with ``|c| <= eps`` (tested as ``|c|^2 > eps^2``)."""
    return Coeff(float(eps))


def weight(k: int) -> Truncation:
    """Drop terms of Pauli weight above ``k``."""
    if k < 0:
        raise ValueError(f"weight cutoff must be non-negative, got {k}")
    return Weight(int(k))


def topn(n: int) -> Truncation:
    """Keep the ``n`` terms of largest ``|c|`` after every layer."""
    if n < 0:
        raise ValueError(f"topn count must be non-negative, got {n}")
    return TopN(int(n))


def approx_topn(n: int) -> Truncation:
    """Keep ``n`` terms after every layer, ranked by octave of ``|c|^2``: whole octaves from the top while they fit,
    the next octave filled in canonical (bucket, then key) order."""
    if n < 0:
        raise ValueError(f"approx_topn count must be non-negative, got {n}")
    return ApproxTopN(int(n))


def lower(policy: Truncation | None) -> "_native.Policy":
    if policy is None:
        return _native.Policy()
    return _check(policy).lower()
