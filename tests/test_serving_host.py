"""Serving-loop request generator (CPU): seeded Poisson arrivals and the
uniform qubit-count, depth and precision draws."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs


def test_poisson_requests_are_seeded_and_in_range():
    a = cs.poisson_requests(3000, 200.0, 12, 28, 10, 50, 0.5, 99)
    b = cs.poisson_requests(3000, 200.0, 12, 28, 10, 50, 0.5, 99)
    assert [(r.arrival_s, r.n_qubits, r.depth, r.seed, r.precision) for r in a] == \
           [(r.arrival_s, r.n_qubits, r.depth, r.seed, r.precision) for r in b]
    t = np.array([r.arrival_s for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # exponential gaps of mean 1/200 s: mean within 5 standard errors
    assert abs(gaps.mean() - 1 / 200) < 5 * (1 / 200) / np.sqrt(len(gaps))
    n = np.array([r.n_qubits for r in a])
    d = np.array([r.depth for r in a])
    assert n.min() == 12 and n.max() == 28 and d.min() == 10 and d.max() == 50
    f32 = np.mean([r.precision == cs.Precision.F32 for r in a])
    assert abs(f32 - 0.5) < 5 * 0.5 / np.sqrt(len(a))
    assert [r.id for r in a] == list(range(3000))
    assert len({r.seed for r in a}) == 3000
    c = cs.poisson_requests(10, 200.0, 12, 28, 10, 50, 0.5, 100)
    assert [r.seed for r in c] != [r.seed for r in a[:10]]


@pytest.mark.parametrize("args", [(5, 0.0, 12, 28, 10, 50, 0.5, 1), (5, 1.0, 1, 28, 10, 50, 0.5, 1),
                                  (5, 1.0, 12, 40, 10, 50, 0.5, 1), (5, 1.0, 12, 28, 0, 50, 0.5, 1),
                                  (5, 1.0, 12, 28, 10, 5, 0.5, 1), (5, 1.0, 12, 28, 10, 50, 1.5, 1)])
def test_poisson_requests_reject_bad_parameters(args):
    with pytest.raises(cs.CastError):
        cs.poisson_requests(*args)
