"""FP64 CPU oracle for the Sequential MPS encoder (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``qmprs_amd``) runs on the GPU through the
``libqmx`` C-ABI and fails loudly when that library is missing.

It restates, step by step in NumPy/SciPy complex128, what the reference
computes for ``Sequential(backend).prepare_state(sv, chi, num_layers=L,
num_sweeps=S)``:

* ``qmprs/synthesis/mps_encoding/base.py:58-106``      -- prepare_state
* ``qmprs/primitives/mps.py:218-249``                    -- from_statevector
* ``qmprs/primitives/mps.py:302-310, 348-459, 538-563``  -- normalise /
  canonicalise / compress / permute
* ``qmprs/primitives/mps.py:565-891``                    -- Ran layer
  extraction (first / two / last site unitaries, bond-2 truncation)
* ``qmprs/primitives/mps.py:933-995``                    -- inverse layer
* ``qmprs/primitives/mps.py:1020-1039``                  -- overlap with |0..0>
* ``qmprs/synthesis/mps_encoding/sequential.py:155-600`` -- layer loop,
  environment sweeps, circuit emission

The reference delegates its arithmetic to quimb 1.10.0 (absent here, see the
``[Q]`` notes below), so two restatements are given:

``layer_mode="mps"``
    quimb-like MPS operations (SVD chain ``from_dense``, QR/LQ canonisation,
    truncated SVD sweeps, ``gate_split_`` with a cutoff-only SVD).  Small N
    only: bonds double per layer.
``layer_mode="dense"``
    the same mathematics on the dense statevector.  Every MPS the reference
    builds is only ever used through the state it represents, except the
    bond-2 MPS whose site tensors become gates; that one is rebuilt from the
    dense state by the right-to-left truncated-projection chain
    (``truncation_chain``).  This is the algorithm the GPU runs.

Gauge.  The Ran gates depend on the bond gauge of the bond-2 MPS.  quimb leaves
it to LAPACK's SVD phases, which no other implementation reproduces.  The
oracle therefore fixes one explicit gauge (``gauge="canonical"``): each
truncated bond is spanned by the Gram-Schmidt orthonormalisation of the
columns of the projector onto the kept subspace, taken in index order.  This
depends only on the subspace, so two correct implementations agree to
rounding.  ``gauge="lapack"`` instead keeps NumPy's SVD vectors, and is used to
measure how much the fidelity depends on the gauge.

Parity pins (tests/test_oracle_*.py): scipy.linalg.null_space, scipy.linalg.polar,
brute-force SVD of reshaped statevectors, Eckart-Young optimality, isometry
residuals, the layer mapping |0..0> onto the bond-2 MPS, dense-vs-MPS
restatement agreement, monotone sweeps, and the reference's integration
scenarios (tests/synthesis/mps_encoding/test_sequential_encoding.py).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

CUTOFF = 1e-10            # quimb default split cutoff [Q]
UNITARY_RTOL = 1e-5       # numpy.allclose defaults, as quick.is_unitary_matrix [Q]
UNITARY_ATOL = 1e-8
GS_TOL = 1e-7             # canonical-gauge Gram-Schmidt skip threshold
PHASE_TOL = 1e-12         # |first entry| below this: phase fix of mps.py:661-662 skipped


# --------------------------------------------------------------------------
# Inputs
# --------------------------------------------------------------------------

def ket(statevector) -> np.ndarray:
    """quick ``Ket``: flatten, cast to complex128 and normalise (base.py:96-97) [Q]."""
    sv = np.asarray(statevector, dtype=np.complex128).reshape(-1)
    n = sv.size
    if n < 2 or (n & (n - 1)) != 0:
        raise ValueError("The statevector length must be a power of two.")
    nrm = np.linalg.norm(sv)
    if nrm == 0:
        raise ValueError("The statevector must be non-zero.")
    return sv / nrm


def random_uniform_state(num_qubits: int, seed: int) -> np.ndarray:
    """Uniform[0,1) real and imaginary parts, normalised (README.md:52-53)."""
    rng = np.random.default_rng(seed)
    sv = rng.random(2 ** num_qubits) + 1j * rng.random(2 ** num_qubits)
    return sv / np.linalg.norm(sv)


def random_mps_state(num_qubits: int, bond: int, seed: int) -> np.ndarray:
    """Statevector of a random MPS with complex Gaussian tensors.

    Bonds are min(bond, 2^i, 2^(N-i)); entries (N(0,1)+iN(0,1))/sqrt(2);
    contracted left to right in complex128 and normalised.
    """
    rng = np.random.default_rng(seed)
    dims = [1] + [min(bond, 2 ** i, 2 ** (num_qubits - i)) for i in range(1, num_qubits)] + [1]
    psi = np.ones((1, 1), dtype=np.complex128)
    for i in range(num_qubits):
        a = (rng.standard_normal((dims[i], 2, dims[i + 1]))
             + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))) / np.sqrt(2.0)
        psi = (psi @ a.reshape(dims[i], 2 * dims[i + 1])).reshape(-1, dims[i + 1])
    psi = psi.reshape(-1)
    return psi / np.linalg.norm(psi)


def change_indexing(statevector, index_type: str) -> np.ndarray:
    """quick ``Ket.change_indexing`` (base.py:99) [Q]: "row" keeps the order;
    "snake" (N >= 3) views the amplitudes as a matrix of two rows and
    reverses the odd row, i.e. the second half of the vector is reversed.
    Parity unpinned: quick is not installed, this states the assumed rule."""
    sv = np.asarray(statevector).reshape(-1).copy()
    if index_type == "row":
        return sv
    if index_type != "snake":
        raise ValueError("Index type not supported.")
    if sv.size >= 8:
        m = sv.reshape(2, -1)
        m[1::2, :] = m[1::2, ::-1]
        sv = m.reshape(-1)
    return sv


def compress_amplitudes(statevector, compression_percentage: float) -> np.ndarray:
    """quick ``Ket.compress`` (base.py:101-102) [Q]: the int(pct / 100 * 2^N)
    amplitudes of smallest magnitude are set to zero; equal magnitudes are
    taken in index order (a stable sort; quick's own tie order is that of
    numpy's default argsort).  Parity unpinned, as ``change_indexing``."""
    sv = np.asarray(statevector).reshape(-1).copy()
    cutoff = int((compression_percentage / 100.0) * sv.size)
    if cutoff <= 0:
        return sv
    mag2 = sv.real.astype(np.float64) ** 2 + sv.imag.astype(np.float64) ** 2
    order = np.argsort(mag2, kind="stable")
    sv[order[:cutoff]] = 0
    return sv


# --------------------------------------------------------------------------
# Truncation rules
# --------------------------------------------------------------------------

def keep_count(sq_values: np.ndarray, max_bond: int | None, cutoff: float = CUTOFF) -> int:
    """Number of singular values kept by quimb's 'rsum2' cutoff and max_bond [Q].

    ``sq_values`` are squared singular values (eigenvalues of the Gram matrix)
    sorted in decreasing order.  The smallest k whose discarded tail weight is
    at most ``cutoff`` times the total is kept, then capped at ``max_bond``;
    at least one value is always kept.
    """
    lam = np.clip(np.asarray(sq_values, dtype=np.float64), 0.0, None)
    total = lam.sum()
    n = lam.size
    k = n
    if total > 0 and cutoff > 0:
        tail = 0.0
        k = n
        for j in range(n - 1, 0, -1):
            tail += lam[j]
            if tail <= cutoff * total:
                k = j
            else:
                break
    if max_bond is not None:
        k = min(k, max_bond)
    return max(k, 1)


def canonical_basis(v_keep: np.ndarray, tol: float = GS_TOL) -> np.ndarray:
    """Canonical orthonormal basis (columns) of span(v_keep).

    Gram-Schmidt (two passes) of the columns of the projector P = V V^H in
    index order; a column whose residual norm is below ``tol`` is skipped.
    The result depends only on the subspace.
    """
    n, k = v_keep.shape
    proj = v_keep @ v_keep.conj().T
    basis = []
    for j in range(n):
        if len(basis) == k:
            break
        v = proj[:, j].copy()
        for _ in range(2):
            for b in basis:
                v = v - b * (b.conj() @ v)
        nv = np.linalg.norm(v)
        if nv > tol:
            basis.append(v / nv)
    if len(basis) < k:  # numerically impossible for a rank-k projector
        raise ValueError("canonical_basis: projector rank below the kept count")
    return np.stack(basis, axis=1)


def dominant_subspace(gram: np.ndarray, max_bond: int | None, cutoff: float,
                      gauge: str) -> tuple[np.ndarray, np.ndarray]:
    """Kept right-singular subspace of Y from its Gram matrix rho = Y^H Y.

    Returns (W, sq_values): W (n x k) with orthonormal columns spanning the
    top-k eigenspace, and all eigenvalues in decreasing order.
    """
    lam, vec = np.linalg.eigh(0.5 * (gram + gram.conj().T))
    order = np.argsort(lam)[::-1]
    lam = lam[order]
    vec = vec[:, order]
    k = keep_count(lam, max_bond, cutoff)
    if gauge == "canonical":
        w = canonical_basis(vec[:, :k])
    elif gauge == "lapack":
        w = vec[:, :k]
    elif gauge.startswith("phases:"):
        # the eigenvectors with random phases: the freedom an SVD leaves on a
        # bond with distinct singular values (measures gauge sensitivity)
        rng = np.random.default_rng([int(gauge.split(":", 1)[1]), n_calls_for_seed(gram)])
        w = vec[:, :k] * np.exp(2j * np.pi * rng.random(k))
    else:
        raise ValueError(f"unknown gauge {gauge!r}")
    return w, lam


def n_calls_for_seed(gram: np.ndarray) -> int:
    """A per-call seed component for the random-phase gauge: derived from the
    matrix itself, so every bond of every layer draws different phases while
    a run stays reproducible."""
    return int(abs(hash(np.round(gram, 6).tobytes())) % (2 ** 31))


# --------------------------------------------------------------------------
# Dense restatement of MPS construction / compression (mps.py:218-249, 402-459)
# --------------------------------------------------------------------------

def truncation_chain(psi: np.ndarray, num_sites: int, max_bond: int | None,
                     cutoff: float = CUTOFF, gauge: str = "canonical",
                     skip_full_rank: bool = False):
    """Right-to-left truncated projection chain on a dense state.

    Equals quimb's ``left_canonize()`` + ``right_compress(max_bond, cutoff)``
    (mps.py:449-453) and its 'dm' ``tensor_network_1d_compress`` (mps.py:247),
    both of which pick at every bond, right to left, the dominant right singular
    subspace of the state already projected on the bonds to its right.

    X_N = psi as (2^N x 1).  For i = N-1 .. 1:
        Y   = X_{i+1} viewed as (2^i x 2 r_{i+1})   [columns (p_i, bond)]
        rho = Y^H Y
        W   = kept subspace of rho (r_i = k columns)
        B_i = W^H  (site tensor, 'lpr': (r_i, 2, r_{i+1}))
        X_i = Y W
    Site 0 is X_1 as (1, 2, r_1) -- the orthogonality centre.

    With ``skip_full_rank`` a step whose column count 2 r_{i+1} does not
    exceed ``max_bond`` keeps the identity basis (no eigensolve, cutoff not
    applied): the represented state is unchanged by such a step.

    Returns (tensors, spectra) with tensors[i] in 'lpr' and spectra[i] the
    eigenvalues of rho at bond (i-1, i) (None where skipped).
    """
    x = psi.reshape(-1, 1).astype(np.complex128)
    tensors = [None] * num_sites
    spectra = [None] * num_sites
    r = 1
    for i in range(num_sites - 1, 0, -1):
        y = x.reshape(2 ** i, 2 * r)
        if skip_full_rank and max_bond is not None and 2 * r <= max_bond:
            w = np.eye(2 * r, dtype=np.complex128)
        else:
            w, lam = dominant_subspace(y.conj().T @ y, max_bond, cutoff, gauge)
            spectra[i] = lam
        k = w.shape[1]
        tensors[i] = w.conj().T.reshape(k, 2, r)
        x = y @ w
        r = k
    tensors[0] = x.reshape(1, 2, r)
    return tensors, spectra


def mps_to_dense(tensors) -> np.ndarray:
    """Contract an 'lpr' MPS (list of (l, 2, r) arrays) to a dense vector."""
    z = tensors[0].reshape(-1, tensors[0].shape[-1])
    for a in tensors[1:]:
        l, _, r = a.shape
        z = (z @ a.reshape(l, 2 * r)).reshape(-1, r)
    return z.reshape(-1)


def chi_truncate(psi: np.ndarray, num_sites: int, bond_dimension: int,
                 cutoff: float = CUTOFF):
    """``MPS.from_statevector`` (mps.py:218-249): dense form of the chi-MPS.

    Returns (target, spectra): the statevector of the chi-truncated MPS (not
    renormalised, as quimb leaves it) and the Gram spectra per bond.
    """
    tensors, spectra = truncation_chain(psi, num_sites, bond_dimension, cutoff,
                                        gauge="lapack", skip_full_rank=True)
    return mps_to_dense(tensors), spectra


# --------------------------------------------------------------------------
# quimb-like MPS operations (layer_mode="mps")  [Q]
# --------------------------------------------------------------------------

def _svd_trunc(m: np.ndarray, max_bond: int | None, cutoff: float):
    u, s, vh = np.linalg.svd(m, full_matrices=False)
    k = keep_count(s ** 2, max_bond, cutoff)
    return u[:, :k], s[:k], vh[:k]


def mps_from_dense(psi: np.ndarray, num_sites: int, cutoff: float = CUTOFF):
    """quimb ``MatrixProductState.from_dense``: right-to-left SVD splits with
    the singular values absorbed into the left remainder (mps.py:242) [Q]."""
    tensors = [None] * num_sites
    rem = psi.reshape(-1, 1)
    r = 1
    for i in range(num_sites - 1, 0, -1):
        m = rem.reshape(2 ** i, 2 * r)
        u, s, vh = _svd_trunc(m, None, cutoff)
        k = s.size
        tensors[i] = vh.reshape(k, 2, r)
        rem = u * s
        r = k
    tensors[0] = rem.reshape(1, 2, r)
    return tensors


def _qr_pos(m: np.ndarray):
    """QR with a real non-negative diagonal of R (quimb 'stabilised' QR) [Q]."""
    q, r = np.linalg.qr(m)
    d = np.diag(r)
    mag = np.abs(d)
    ph = np.ones_like(d)
    nz = mag > 0
    ph[nz] = d[nz] / mag[nz]
    return q * ph, ph.conj()[:, None] * r


def mps_left_canonize(tensors):
    t = [a.copy() for a in tensors]
    for i in range(len(t) - 1):
        l, _, r = t[i].shape
        q, rr = _qr_pos(t[i].reshape(l * 2, r))
        k = q.shape[1]
        t[i] = q.reshape(l, 2, k)
        l2, _, r2 = t[i + 1].shape
        t[i + 1] = (rr @ t[i + 1].reshape(l2, 2 * r2)).reshape(k, 2, r2)
    return t


def mps_right_canonize(tensors, normalize: bool = False):
    t = [a.copy() for a in tensors]
    for i in range(len(t) - 1, 0, -1):
        l, _, r = t[i].shape
        q, rr = _qr_pos(t[i].reshape(l, 2 * r).T)
        k = q.shape[1]
        t[i] = q.T.reshape(k, 2, r)
        lp, _, rp = t[i - 1].shape
        t[i - 1] = (t[i - 1].reshape(lp * 2, rp) @ rr.T).reshape(lp, 2, k)
    if normalize:
        t[0] = t[0] / np.linalg.norm(t[0])
    return t


def mps_right_compress(tensors, max_bond: int | None, cutoff: float = CUTOFF,
                       gauge: str = "lapack"):
    """Right-to-left SVD truncation sweep from a left-canonical MPS [Q]."""
    t = [a.copy() for a in tensors]
    for i in range(len(t) - 1, 0, -1):
        l, _, r = t[i].shape
        m = t[i].reshape(l, 2 * r)
        w, _ = dominant_subspace(m.conj().T @ m, max_bond, cutoff, gauge)
        k = w.shape[1]
        t[i] = w.conj().T.reshape(k, 2, r)
        lp, _, rp = t[i - 1].shape
        t[i - 1] = (t[i - 1].reshape(lp * 2, rp) @ (m @ w)).reshape(lp, 2, k)
    return t


def mps_gate_single(tensors, gate: np.ndarray, site: int):
    """quimb ``gate_(G, where=[site], contract=True)`` (mps.py:953-957)."""
    t = list(tensors)
    t[site] = np.einsum("op,lpr->lor", gate, t[site])
    return t


def mps_gate_split(tensors, gate: np.ndarray, site: int, cutoff: float = CUTOFF):
    """quimb ``gate_split_(G, where=[site, site+1])`` (mps.py:968-971) [Q]:
    theta = G . (A_i A_{i+1}), SVD with cutoff only, sqrt(S) absorbed both ways."""
    t = list(tensors)
    a, b = t[site], t[site + 1]
    l, _, m = a.shape
    _, _, r = b.shape
    theta = np.einsum("lpm,mqr->lpqr", a, b).reshape(l, 4, r)
    theta = np.einsum("xy,lyr->lxr", gate, theta).reshape(l * 2, 2 * r)
    u, s, vh = _svd_trunc(theta, None, cutoff)
    k = s.size
    sq = np.sqrt(s)
    t[site] = (u * sq).reshape(l, 2, k)
    t[site + 1] = (sq[:, None] * vh).reshape(k, 2, r)
    return t


# --------------------------------------------------------------------------
# Dense gate application (MSB convention: site 0 is the most significant bit)
# --------------------------------------------------------------------------

def apply_gate(psi: np.ndarray, num_sites: int, gate: np.ndarray, site: int,
               width: int) -> np.ndarray:
    """Apply a (2^w x 2^w) gate, (out) x (in), to sites site..site+w-1."""
    d = 2 ** width
    v = psi.reshape(2 ** site, d, 2 ** (num_sites - site - width))
    return np.einsum("oi,aib->aob", gate, v, optimize=True).reshape(-1)


def environment(tp: np.ndarray, phi: np.ndarray, num_sites: int, site: int,
                width: int) -> np.ndarray:
    """E[o, i] = sum_rest conj(T'[o, rest]) * phi[i, rest] (sequential.py:463)."""
    d = 2 ** width
    a = tp.reshape(2 ** site, d, -1)
    b = phi.reshape(2 ** site, d, -1)
    return np.einsum("aob,aib->oi", a.conj(), b, optimize=True)


# --------------------------------------------------------------------------
# Ran layer extraction (mps.py:565-891)
# --------------------------------------------------------------------------

def householder_null_space(rows: np.ndarray) -> np.ndarray:
    """Null space of ``conj(rows)``, i.e. vectors orthogonal to the rows.

    Restates ``scipy.linalg.null_space(rows.conj())`` (mps.py:597,657,730) for
    inputs with orthonormal rows as the trailing columns of the full
    Householder QR of rows^T (LAPACK zgeqr2 + zung2r conventions, zlarfg
    reflectors with beta = -sign(Re alpha) * ||x||).
    """
    a = rows.T.astype(np.complex128).copy()       # (n x m) = (conj(rows))^H
    n, m = a.shape
    taus = []
    vs = []
    for j in range(m):
        alpha = a[j, j]
        x = a[j + 1:, j]
        xnorm = np.linalg.norm(x)
        if xnorm == 0.0 and alpha.imag == 0.0:
            tau = 0.0 + 0.0j
            v = np.zeros(n - j, dtype=np.complex128)
            v[0] = 1.0
        else:
            beta = -np.copysign(np.sqrt(alpha.real ** 2 + alpha.imag ** 2 + xnorm ** 2), alpha.real)
            tau = complex((beta - alpha.real) / beta, -alpha.imag / beta)
            scale = 1.0 / (alpha - beta)
            v = np.concatenate(([1.0 + 0j], x * scale))
            a[j, j] = beta
            a[j + 1:, j] = 0
        # apply H^H = I - conj(tau) v v^H to the trailing columns
        if j + 1 < m:
            sub = a[j:, j + 1:]
            sub -= np.conj(tau) * np.outer(v, v.conj() @ sub)
        taus.append(tau)
        vs.append(v)
    q = np.eye(n, dtype=np.complex128)
    for j in range(m - 1, -1, -1):
        v, tau = vs[j], taus[j]
        sub = q[j:, :]
        sub -= tau * np.outer(v, v.conj() @ sub)
    return q[:, m:]


def submps_indices(bonds) -> list[tuple[int, int]]:
    """quick ``_get_submps_indices`` (mps.py:794-802) [Q]: maximal runs of sites
    joined by bonds of dimension >= 2; isolated sites form (i, i) blocks."""
    n = len(bonds) + 1
    blocks = []
    start = 0
    for i in range(n):
        right = bonds[i] if i < n - 1 else 1
        if right < 2:
            blocks.append((start, i))
            start = i + 1
    return blocks


def is_unitary(m: np.ndarray) -> bool:
    return bool(np.allclose(m @ m.conj().T, np.eye(m.shape[0]),
                            rtol=UNITARY_RTOL, atol=UNITARY_ATOL))


def first_site_unitary(a: np.ndarray) -> np.ndarray:
    """mps.py:565-623 (Ran eq. 9): M = [a | n1 | n0 | n2], a = A.flatten()."""
    v = a.reshape(-1)
    nb = householder_null_space(v.reshape(1, -1))
    return np.stack([v, nb[:, 1], nb[:, 0], nb[:, 2]], axis=1)


def two_site_unitary(a: np.ndarray) -> np.ndarray:
    """mps.py:625-687 (Ran eq. 7): M = [A[0] | n0 | A[1] | n1] with the null
    basis columns rotated so their first entries are real (mps.py:661-662).
    A first entry below PHASE_TOL in magnitude has a rounding-determined phase
    in any implementation; such a column is left unrotated."""
    b = a.reshape(2, 4)
    nb = householder_null_space(b)
    # a first entry at rounding level has no meaningful phase: no rotation
    first = nb[0]
    rot = np.where(np.abs(first) > PHASE_TOL, np.exp(1j * np.angle(first)), 1.0)
    nb = nb / rot
    return np.stack([b[0], nb[:, 0], b[1], nb[:, 1]], axis=1)


def last_site_unitary(a: np.ndarray, single: bool) -> np.ndarray:
    """mps.py:689-744 (Ran eq. 6): M[p, l] = A[l, p]; isolated site: rows
    [a; null(conj a)] transposed."""
    if single:
        v = a.reshape(-1)
        nb = householder_null_space(v.reshape(1, -1))
        u = np.stack([v, nb[:, 0]], axis=0)
        return u.T
    return a.reshape(2, 2).T


def generate_layer(tensors):
    """``generate_unitary_layer`` (mps.py:746-847) on a bond-2 'lpr' MPS.

    Returns a list of blocks (start, end, [gates]) with gates (out x in).
    """
    n = len(tensors)
    bonds = [tensors[i].shape[2] for i in range(n - 1)]
    layer = []
    for start, end in submps_indices(bonds):
        gates = []
        for i in range(start, end + 1):
            if i == end:
                g = last_site_unitary(tensors[i], start == end)
            elif i == start:
                g = first_site_unitary(tensors[i])
            else:
                g = two_site_unitary(tensors[i])
            gates.append(g)
        for g in gates:
            if not is_unitary(g):
                raise ValueError("All the generated unitaries must be unitary.")
        layer.append((start, end, gates))
    return layer


def bond2_mps(psi_w: np.ndarray, num_sites: int, gauge: str = "canonical",
              cutoff: float = CUTOFF):
    """``generate_bond_D_unitary_layer`` preamble (mps.py:876-887): compress to
    bond 2 in right form, then right-canonise with normalisation."""
    tensors, _ = truncation_chain(psi_w, num_sites, 2, cutoff, gauge=gauge)
    tensors[0] = tensors[0] / np.linalg.norm(tensors[0])
    return tensors


def apply_inverse_layer_dense(psi: np.ndarray, num_sites: int, layer) -> np.ndarray:
    """``_apply_inverse_unitary_layer`` (mps.py:933-971) on a dense state."""
    for start, end, gates in layer:
        for i in range(end, start - 1, -1):
            g = gates[i - start].conj().T
            psi = apply_gate(psi, num_sites, g, i, 1 if i == end else 2)
    return psi


def apply_inverse_layer_mps(tensors, layer, cutoff: float = CUTOFF):
    for start, end, gates in layer:
        for i in range(end, start - 1, -1):
            g = gates[i - start].conj().T
            if i == end:
                tensors = mps_gate_single(tensors, g, i)
            else:
                tensors = mps_gate_split(tensors, g, i, cutoff)
    return tensors


def layer_gate_list(layers):
    """Gates in circuit order as (site, width, matrix), following
    ``_tensor_network_from_unitary_layers`` (sequential.py:244-270)."""
    out = []
    for layer in layers:
        for start, end, gates in layer:
            for i in range(start, end + 1):
                out.append((i, 1 if i == end else 2, gates[i - start]))
    return out


def get_unitary_layers(target: np.ndarray, num_sites: int, num_layers: int,
                       fidelity_threshold: float = 1 - 1e-6, gauge: str = "canonical",
                       layer_mode: str = "dense", cutoff: float = CUTOFF,
                       return_info: bool = False):
    """``Sequential._get_unitary_layers`` (sequential.py:330-398)."""
    info = {"amp0": [], "bonds": []}
    if layer_mode == "dense":
        psi_w = target / np.linalg.norm(target)
    elif layer_mode == "mps":
        t = mps_from_dense(target, num_sites, cutoff)
        t = mps_left_canonize(t)
        t = mps_right_compress(t, None, cutoff)          # compress(mode="right")
        t = mps_right_canonize(t, normalize=True)        # canonicalize("right", True)
    else:
        raise ValueError(layer_mode)
    layers = []
    for _ in range(num_layers):
        if layer_mode == "dense":
            b2 = bond2_mps(psi_w, num_sites, gauge, cutoff)
        else:
            c = mps_left_canonize(t)
            c = mps_right_compress(c, 2, cutoff, gauge)
            b2 = mps_right_canonize(c, normalize=True)
        layer = generate_layer(b2)
        layers.append(layer)
        if layer_mode == "dense":
            psi_w = apply_inverse_layer_dense(psi_w, num_sites, layer)
            amp0 = np.conj(psi_w[0])
        else:
            t = apply_inverse_layer_mps(t, layer, cutoff)
            info["bonds"].append([t[i].shape[2] for i in range(num_sites - 1)])
            amp0 = np.conj(mps_to_dense(t)[0])
        info["amp0"].append(amp0)
        if np.isclose(amp0, 1 + 0j, atol=1 - fidelity_threshold):
            break
    layers.reverse()
    if return_info:
        return layers, info
    return layers


# --------------------------------------------------------------------------
# Environment sweeps (sequential.py:400-541)
# --------------------------------------------------------------------------

def polar_unitary(env: np.ndarray) -> np.ndarray:
    """u @ vh of the SVD of E (sequential.py:473-478)."""
    u, _, vh = np.linalg.svd(env)
    return u @ vh


def _complete(basis: list, d: int, count: int) -> list:
    """Extend orthonormal columns by two-pass Gram-Schmidt of e_0, e_1, ..."""
    out = []
    for e0 in range(d):
        if len(out) == count:
            break
        x = np.zeros(d, dtype=np.complex128)
        x[e0] = 1.0
        for _ in range(2):
            for b in basis + out:
                x = x - b * np.vdot(b, x)
        nx = np.linalg.norm(x)
        if nx > 0.5:
            out.append(x / nx)
    return out


def polar_unitary_jacobi(env: np.ndarray, tol: float = 1e-13) -> np.ndarray:
    """The same polar factor by a one-sided (Hestenes) Jacobi SVD.

    For a rank-deficient E the unitary polar factor is free on the null space
    of E; LAPACK fills it with whatever singular vectors gesdd returns.  This
    variant fixes that part canonically: with r singular values above ``tol``
    * sigma_max, the polar factor is sum_j u_j v_j^H over those r pairs plus
    sum_c x_c y_c^H, where x_c and y_c complete {u_j} and {v_j} by two-pass
    Gram-Schmidt of the unit vectors e_0, e_1, ... in order.  The result
    depends only on the well-determined singular subspaces.
    Rotations: cyclic pairs (p < q), rotate when |a_p^H a_q| > 1e-16 |a_p||a_q|.
    """
    d = env.shape[0]
    a = env.astype(np.complex128).copy()
    v = np.eye(d, dtype=np.complex128)
    for _ in range(60):
        rotated = False
        for p in range(d - 1):
            for q in range(p + 1, d):
                al = np.vdot(a[:, p], a[:, p]).real
                be = np.vdot(a[:, q], a[:, q]).real
                g = np.vdot(a[:, p], a[:, q])
                ag = abs(g)
                if ag == 0.0 or ag <= 1e-16 * np.sqrt(al * be):
                    continue
                rotated = True
                ph = g / ag
                zeta = (be - al) / (2.0 * ag)
                t = (1.0 if zeta >= 0 else -1.0) / (abs(zeta) + np.sqrt(1.0 + zeta * zeta))
                c = 1.0 / np.sqrt(1.0 + t * t)
                s = t * c
                rot = np.array([[c, s * ph], [-s * np.conj(ph), c]])
                a[:, [p, q]] = a[:, [p, q]] @ rot
                v[:, [p, q]] = v[:, [p, q]] @ rot
        if not rotated:
            break
    sig = np.linalg.norm(a, axis=0)
    smax = sig.max()
    good = [j for j in range(d) if smax > 0 and sig[j] > tol * smax]
    us = [a[:, j] / sig[j] for j in good]
    vs = [v[:, j] for j in good]
    out = np.zeros((d, d), dtype=np.complex128)
    for u_, v_ in zip(us, vs):
        out += np.outer(u_, v_.conj())
    for x, y in zip(_complete(us, d, d - len(us)), _complete(vs, d, d - len(vs))):
        out += np.outer(x, y.conj())
    return out


def circuit_state(num_sites: int, gate_list) -> np.ndarray:
    psi = np.zeros(2 ** num_sites, dtype=np.complex128)
    psi[0] = 1.0
    for site, width, g in gate_list:
        psi = apply_gate(psi, num_sites, g, site, width)
    return psi


def relaxed_update(g_new: np.ndarray, g_old: np.ndarray, omega: float) -> np.ndarray:
    """Over-relaxed gate update (no reference counterpart; the reference's
    update is omega = 1): polar(G' + (omega - 1)(G' - G)), with G' the plain
    update conj(polar(E)) and G the gate before it."""
    if omega == 1.0:
        return g_new
    return polar_unitary(g_new + (omega - 1.0) * (g_new - g_old))


def sweep_once(target: np.ndarray, num_sites: int, gate_list, polar: str = "lapack",
               omega: float = 1.0):
    """One ``_sweep_unitary_layers`` pass (sequential.py:400-507).

    phi <- G^dag phi; E = <T'| phi>_rest; G' = conj(polar(E)); T' <- G'^dag T'.
    Returns the updated gate list and the overlap <T|C'|0> after the last
    update, sum_{o,i} G'[o,i] E[o,i].  ``omega`` != 1 replaces G' by
    ``relaxed_update`` (not part of the reference).
    """
    gl = list(gate_list)
    phi = circuit_state(num_sites, gl)
    tp = target.astype(np.complex128).copy()
    overlap = 0j
    for idx in range(len(gl) - 1, -1, -1):
        site, width, g = gl[idx]
        phi = apply_gate(phi, num_sites, g.conj().T, site, width)
        env = environment(tp, phi, num_sites, site, width)
        pu = polar_unitary(env) if polar == "lapack" else polar_unitary_jacobi(env)
        g_new = relaxed_update(np.conj(pu), g, omega)
        tp = apply_gate(tp, num_sites, g_new.conj().T, site, width)
        gl[idx] = (site, width, g_new)
        overlap = np.sum(g_new * env)
    return gl, overlap


def optimize_gate_list(target, num_sites, gate_list, num_sweeps, trace=None, polar="lapack",
                       omega: float = 1.0):
    """``_optimize_unitary_layers`` (sequential.py:509-541)."""
    for _ in range(num_sweeps):
        gate_list, ov = sweep_once(target, num_sites, gate_list, polar, omega)
        if trace is not None:
            trace.append(ov)
    return gate_list


def gate_list_to_layers(gate_list, layers_template):
    """Write an updated flat gate list back into the layer structure
    (sequential.py:498-505)."""
    it = iter(gate_list)
    out = []
    for layer in layers_template:
        nl = []
        for start, end, gates in layer:
            ng = [next(it)[2] for _ in gates]
            nl.append((start, end, ng))
        out.append(nl)
    return out


# --------------------------------------------------------------------------
# End to end (base.py:58-106 -> sequential.py:543-600)
# --------------------------------------------------------------------------

def prepare_state(statevector, bond_dimension: int, num_layers: int = 1,
                  num_sweeps: int = 0, fidelity_threshold: float = 1 - 1e-6,
                  gauge: str = "canonical", layer_mode: str = "dense",
                  cutoff: float = CUTOFF, polar: str = "lapack"):
    """Returns a dict with the unitary layers (circuit order), the flat gate
    list, the dense circuit state C|0..0>, the fidelity |<psi|C|0>| and the
    chi-truncated target.  ``polar``: "lapack" (scipy) or "jacobi" (the
    canonical completion of a rank-deficient environment)."""
    if not isinstance(num_layers, int) or num_layers < 1:
        raise ValueError("The number of layers must be a positive integer.")
    if not isinstance(bond_dimension, int) or bond_dimension < 1:
        raise ValueError("`bond_dimension` must be an integer greater than 0.")
    psi = ket(statevector)
    n = int(np.log2(psi.size))
    if n < 2:
        raise ValueError("The statevector must have at least 2 qubits.")
    target, spectra = chi_truncate(psi, n, bond_dimension, cutoff)
    layers = get_unitary_layers(target, n, num_layers, fidelity_threshold, gauge,
                                layer_mode, cutoff)
    gl = layer_gate_list(layers)
    trace = []
    if num_sweeps > 0:
        gl = optimize_gate_list(target, n, gl, num_sweeps, trace, polar)
        layers = gate_list_to_layers(gl, layers)
    out_state = circuit_state(n, gl)
    fid = abs(np.vdot(psi, out_state))
    return {"layers": layers, "gates": gl, "state": out_state, "fidelity": fid,
            "target": target, "spectra": spectra, "num_qubits": n,
            "overlap_trace": trace, "psi": psi}


# --------------------------------------------------------------------------
# Alternative schedules (notebook "State Preparation using MPS Sequential",
# markdown cell 22; Rudolph et al., arXiv:2209.00595).  The reference runs
# DallOall (all layers decomposed, then all optimised: prepare_state above)
# and names two iterative variants without implementing them:
#   IterDiOi   -- decompose layer i from the current residual, optimise that
#                 layer alone against the residual, then peel it off;
#   IterDiOall -- decompose layer i, optimise every layer so far against the
#                 chi-MPS target, and recompute the residual C^dag |target>.
# Parity unpinned against the reference (it has no code for them); pinned by
# their reductions to DallOall (num_sweeps = 0, or num_layers = 1) and by the
# monotone local updates of the sweep.
# --------------------------------------------------------------------------

SCHEDULES = ("DallOall", "IterDiOi", "IterDiOall")


def schedule_layers(target: np.ndarray, num_sites: int, num_layers: int, num_sweeps: int,
                    schedule: str, fidelity_threshold: float = 1 - 1e-6,
                    cutoff: float = CUTOFF, polar: str = "lapack"):
    """Layers (circuit order) of an iterative schedule on the chi-MPS state
    ``target``; the early exit of sequential.py:384-391 is checked on the
    residual after each layer."""
    if schedule not in ("IterDiOi", "IterDiOall"):
        raise ValueError(f"unknown schedule {schedule!r}")
    unit = target / np.linalg.norm(target)
    resid = unit.copy()
    layers = []                      # extraction order
    for _ in range(num_layers):
        layer = generate_layer(bond2_mps(resid, num_sites, "canonical", cutoff))
        if schedule == "IterDiOi":
            gl = optimize_gate_list(resid, num_sites, layer_gate_list([layer]), num_sweeps,
                                    polar=polar)
            layer = gate_list_to_layers(gl, [layer])[0]
            layers.append(layer)
            resid = apply_inverse_layer_dense(resid, num_sites, layer)
        else:
            layers.append(layer)
            circ = layers[::-1]
            gl = optimize_gate_list(target, num_sites, layer_gate_list(circ), num_sweeps,
                                    polar=polar)
            layers = gate_list_to_layers(gl, circ)[::-1]
            resid = unit.copy()
            for lay in layers:
                resid = apply_inverse_layer_dense(resid, num_sites, lay)
        if np.isclose(np.conj(resid[0]), 1 + 0j, atol=1 - fidelity_threshold):
            break
    return layers[::-1]


def prepare_state_schedule(statevector, bond_dimension: int, num_layers: int = 1,
                           num_sweeps: int = 0, schedule: str = "DallOall",
                           fidelity_threshold: float = 1 - 1e-6, cutoff: float = CUTOFF,
                           polar: str = "lapack"):
    """prepare_state under one of SCHEDULES; same result dictionary keys as
    ``prepare_state`` (layers, gates, state, fidelity, target)."""
    if schedule == "DallOall":
        return prepare_state(statevector, bond_dimension, num_layers, num_sweeps,
                             fidelity_threshold, cutoff=cutoff, polar=polar)
    psi = ket(statevector)
    n = int(np.log2(psi.size))
    target, _ = chi_truncate(psi, n, bond_dimension, cutoff)
    layers = schedule_layers(target, n, num_layers, num_sweeps, schedule, fidelity_threshold,
                             cutoff, polar)
    gl = layer_gate_list(layers)
    state = circuit_state(n, gl)
    return {"layers": layers, "gates": gl, "state": state, "fidelity": abs(np.vdot(psi, state)),
            "target": target, "num_qubits": n, "psi": psi}
