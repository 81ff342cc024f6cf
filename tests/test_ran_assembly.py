"""Ran gate assembly: the oracle against the reference's own index algebra.

The reference builds each gate by filling a (2, 2, 2, 2) array slot by slot,
transposing its first two axes, reshaping to 4 x 4 and transposing
(mps.py:594-619 first site, :654-683 two sites, :727-740 last site), with the
null space from ``scipy.linalg.null_space`` and, for two sites, the column
phase fix of :661-662.  ``_ref_*`` below redo exactly those array operations
(numpy + scipy, written here for the test) on random right-canonical site
tensors and compare with ``oracle.first/two/last_site_unitary``, which state
the result in closed form ([a|n1|n0|n2], [A0|n0|A1|n1], M[p, l] = A[l, p]) and
which the GPU's ran_gates_kernel is checked against.  A swapped or
mis-phased column fails these comparisons (checked below by planting such a
mistake).  The one intended difference, null-basis columns whose first entry
is at rounding level (left unrotated by the oracle and the GPU, rotated by a
rounding-noise phase in the reference), is pinned separately: equal up to a
unit phase.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

from oracle import qmprs_oracle as O


def _ref_first(site: np.ndarray) -> np.ndarray:
    """First site of a block: site tensor (p, r)."""
    u = np.zeros((2, 2, 2, 2), dtype=np.complex128)
    basis = scipy.linalg.null_space(site.reshape(1, -1).conj())      # 4 x 3
    u[0, 0] = site.reshape(2, -1)
    u[0, 1] = basis[:, 0].reshape(2, 2)
    u[1, 0] = basis[:, 1].reshape(2, 2)
    u[1, 1] = basis[:, 2].reshape(2, 2)
    u = u.transpose(1, 0, 2, 3)
    return u.reshape(4, 4).T


def _ref_two(site: np.ndarray, phase_fix: bool = True) -> np.ndarray:
    """Inner site of a block: site tensor (l, p, r)."""
    u = np.zeros((2, 2, 2, 2), dtype=np.complex128)
    basis = scipy.linalg.null_space(site.reshape(2, -1).conj())      # 4 x 2
    if phase_fix:
        basis = basis / np.exp(1j * np.angle(basis[0]))
    u[0] = site
    # (p, r, k, 1) -> (1, k, p, r): numpy drops the leading unit axis on assignment
    u[1] = basis.reshape(2, 2, 2, 1).transpose(3, 2, 0, 1)
    u = u.transpose(1, 0, 2, 3)
    return u.reshape(4, 4).T


def _ref_last(site: np.ndarray, single: bool) -> np.ndarray:
    """Last site of a block: (l, p), or an isolated site (p,)."""
    if single:
        u = np.zeros((2, 2), dtype=np.complex128)
        u[0] = site.reshape(1, -1)
        u[1] = scipy.linalg.null_space(site.reshape(1, -1).conj()).reshape(1, -1)
    else:
        u = site
    return u.reshape(2, 2).T


def _unitary(rng, d):
    z = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _random_sites(rng):
    a1 = _unitary(rng, 4)[0]                           # unit 4-vector (p, r)
    a2 = _unitary(rng, 4)[:2].reshape(2, 2, 2)         # orthonormal rows (l; p r)
    a3 = _unitary(rng, 2)                              # (l, p), unitary
    a4 = _unitary(rng, 2)[0]                           # isolated site (p,)
    return a1, a2, a3, a4


def test_oracle_gates_equal_the_reference_index_algebra():
    rng = np.random.default_rng(2024)
    worst = 0.0
    for _ in range(300):
        a1, a2, a3, a4 = _random_sites(rng)
        pairs = [(_ref_first(a1.reshape(2, 2)), O.first_site_unitary(a1.reshape(2, 2))),
                 (_ref_two(a2), O.two_site_unitary(a2)),
                 (_ref_last(a3, False), O.last_site_unitary(a3, False)),
                 (_ref_last(a4, True), O.last_site_unitary(a4, True))]
        for ref, got in pairs:
            d = ref.shape[0]
            assert np.abs(ref @ ref.conj().T - np.eye(d)).max() < 1e-12
            worst = max(worst, np.abs(ref - got).max())
    assert worst < 1e-12, worst


def test_columns_hold_the_site_tensor_where_the_reference_puts_it():
    """Closed forms read off the algebra: the site tensor sits in column 0 of
    the first-site gate, columns 0 and 2 of the two-site gate (rows l = 0, 1),
    and the last-site gate is the transposed tensor."""
    rng = np.random.default_rng(5)
    a1, a2, a3, _ = _random_sites(rng)
    np.testing.assert_array_equal(_ref_first(a1.reshape(2, 2))[:, 0], a1)
    m2 = _ref_two(a2)
    np.testing.assert_array_equal(m2[:, 0], a2[0].reshape(-1))
    np.testing.assert_array_equal(m2[:, 2], a2[1].reshape(-1))
    np.testing.assert_array_equal(_ref_last(a3, False), a3.T)


def test_planted_column_mistakes_are_caught():
    """The comparison is sensitive: the oracle's first-site gate with its two
    null columns swapped, or its two-site gate without the phase fix, differs
    from the reference algebra by O(1)."""
    rng = np.random.default_rng(9)
    a1, a2, _, _ = _random_sites(rng)
    good = O.first_site_unitary(a1.reshape(2, 2))
    swapped = good[:, [0, 2, 1, 3]]
    assert np.abs(swapped - _ref_first(a1.reshape(2, 2))).max() > 0.1
    no_fix = _ref_two(a2, phase_fix=False)
    assert np.abs(no_fix - O.two_site_unitary(a2)).max() > 0.1


def test_rounding_level_first_entries_differ_only_by_a_phase():
    """A site tensor whose row space contains e_0 has null vectors with first
    entry at rounding level: the reference rotates them by the phase of that
    noise, the oracle (and the GPU) leave them unrotated (first entry below
    1e-12).  Both gates are unitary, agree on the site-tensor columns and
    differ on each null column by a unit phase only."""
    rng = np.random.default_rng(13)
    for _ in range(50):
        w = _unitary(rng, 3)[0]
        row0 = np.zeros(4, dtype=np.complex128)
        row0[0] = np.exp(1j * rng.uniform(0, 2 * np.pi))
        row1 = np.concatenate(([0.0], w))
        a2 = np.stack([row0, row1]).reshape(2, 2, 2)
        ref, got = _ref_two(a2), O.two_site_unitary(a2)
        assert np.abs(got @ got.conj().T - np.eye(4)).max() < 1e-12
        np.testing.assert_allclose(got[:, [0, 2]], ref[:, [0, 2]], atol=1e-15)
        for c in (1, 3):
            overlap = np.vdot(ref[:, c], got[:, c])
            assert abs(abs(overlap) - 1.0) < 1e-12
            assert np.abs(got[:, c] - overlap * ref[:, c]).max() < 1e-12
