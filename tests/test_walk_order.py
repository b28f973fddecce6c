"""CPU checks of the pruned eigs walk's order: a time-reversal mirror takes its partner's levels
whatever order the momenta were listed in (mocked blocks; no device)."""

import numpy as np
import pytest

from qed_spin_amd import eigs as E


class _Sec:
    def __init__(self, dim):
        self.dim = dim


class _Mat:
    def __init__(self, est, levels):
        self._est = est
        self.levels = levels

    def estimate(self):
        return self._est


def _walk(blocks, partners, k, monkeypatch):
    monkeypatch.setattr(E, "parallel_build", lambda fn, items: [fn(x) for x in items])
    solved = []

    def solve(m, dim, want, floor, max_iter):
        solved.append(m)
        return np.asarray(m.levels[:want], float), True, "mock"

    monkeypatch.setattr(E, "_solve_block", solve)
    levels, pruned = E.walk_levels(blocks, k, floor=0, partners=partners)
    found = np.sort(np.concatenate([v for v, _, _ in levels.values()]))[:k]
    return found, pruned, solved


def _blocks(order):
    """Block 'p' (momentum 1) and its mirror 'm' (momentum N - 1) share estimate -3 and levels;
    'a' and 'b' lie above. order: the labels in the order the walk lists them."""
    spec = {"p": (-3.0, [-3.0, -2.0]), "m": (-3.0, [-3.0, -2.0]), "a": (-2.5, [-2.5, -1.5]), "b": (-1.0, [-1.0, 0.0])}
    blocks = [(lab, _Sec(100), _Mat(*spec[lab])) for lab in order]
    partners = {order.index("m"): order.index("p")}
    return blocks, partners


@pytest.mark.parametrize("order", [["p", "m", "a", "b"], ["m", "p", "a", "b"], ["a", "m", "b", "p"],
                                   ["b", "a", "p", "m"]])
def test_mirror_takes_partner_levels_in_any_order(order, monkeypatch):
    blocks, partners = _blocks(order)
    found, pruned, solved = _walk(blocks, partners, 4, monkeypatch)
    assert np.array_equal(found, [-3.0, -3.0, -2.5, -2.0]), (order, found)
    assert pruned == 1                      # 'b' only
    assert all(m is not blocks[order.index("m")][2] for m in solved)   # the mirror is never solved


def test_walk_order_keeps_mirror_after_partner():
    # mirror at index 0 of its partner at index 3, an unrelated block of equal estimate between
    est = [-3.0, -1.0, -3.0, -3.0]
    cand = [0, 1, 2, 3]
    order = E.walk_order(est, cand, {0: 3})
    assert order == [2, 3, 0, 1]


def test_mirror_of_pruned_partner_is_pruned(monkeypatch):
    spec = [("a", -3.0, [-3.0, -2.9, -2.8, -2.7]), ("m", 5.0, [5.0]), ("p", 5.0, [5.0])]
    blocks = [(lab, _Sec(100), _Mat(e, lv)) for lab, e, lv in spec]
    found, pruned, _ = _walk(blocks, {1: 2}, 4, monkeypatch)
    assert np.array_equal(found, [-3.0, -2.9, -2.8, -2.7])
    assert pruned == 2
