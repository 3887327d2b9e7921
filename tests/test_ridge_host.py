"""The ridge-tracking oracle (no GPU): the reference's known answer, and
optimality of the tracked path against every path of tiny maps."""
import itertools

import numpy as np

from oracle import ssq_oracle as O


def test_oracle_known_answer():
    """ridge_extraction_test.py:17-26."""
    T = np.array([[1, 4, 4], [2, 2, 2], [5, 5, 4]])
    r, _ = O.ridge_track(T, np.exp([1, 2, 3]), penalty=2.0)
    assert np.array_equal(r[:, 0], [2, 2, 2])


def test_oracle_path_is_least_cost():
    """The forward-backward path minimizes sum_t E[r_t, t] + sum_t
    penalty (s[r_t] - s[r_t-1])^2 (to float32 rounding of the sums)."""
    rng = np.random.default_rng(1)
    for trial in range(6):
        nf, N = 4, 6
        T = (rng.standard_normal((nf, N)) + 1j * rng.standard_normal((nf, N))
             ).astype(np.complex64)
        scales = np.sort(rng.uniform(1, 8, nf)).astype(np.float32)
        pen = [0.5, 2.0][trial % 2]
        r, _ = O.ridge_track(T, scales, penalty=pen)
        energy = np.abs(T.astype(np.complex128))**2
        E = -np.log(energy / energy.max(axis=0) + np.finfo(np.float32).eps)
        s = np.log(scales.astype(np.float64))

        def cost(path):
            return (sum(E[path[t], t] for t in range(N)) +
                    sum(pen * (s[path[t]] - s[path[t - 1]])**2 for t in range(1, N)))
        best = min(cost(p) for p in itertools.product(range(nf), repeat=N))
        assert cost(r[:, 0]) <= best + 1e-4 * (1 + abs(best))
