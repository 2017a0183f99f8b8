"""FP64 CPU oracle for the contraction & truncation hot path.

THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it, and only as the checker (or the
timed CPU baseline) -- never as the product path.  The product is libqtn
(qaintensor.jl_amd/csrc) driven through the qaintensor host package.

It restates, in numpy on LAPACK (zgesdd through numpy.linalg.svd, the same
driver Julia's default svd calls), the semantics of Qaintensor.jl:
  - truncation rule            src/svd.jl:29-33
  - contract_svd               src/svd.jl:7-38
  - ncon labelling of contract src/contract.jl:39-60, 242-264
  - MPS(psi) SVD sweep         src/mps.jl:55-89
  - switch! two-site split     src/switch.jl:18-56
  - MPO(m) / decompose!        src/mpo.jl:27-90, src/decompose.jl:6-52
  - extend_MPO / apply_MPO     src/mpo.jl:122-157, 184-252
and the build's new operations that the reference only has as primitives
(two-site gate application with cutoff + chi_max truncation, amplitudes,
<Z>, overlaps), in the right-canonical Hastings gauge used on the GPU.

Index convention: arrays use the reference's (Julia) index order; Julia's
column-major reshape is reproduced with order='F'.  Wire w <-> bit w-1 of
the state index (least significant first); 2-qubit gate matrices act on
index s_w + 2 s_{w+1} (src/mpo.jl:199-214: apply_MPO contracts the state
with the gate's column legs).

Pinning status of each function is listed in PINNING below; every entry is
exercised by tests/test_oracle.py.
"""
import numpy as np

PINNING = {
    "truncation_rank": "KAT s=exp(-(0:99)), er=1e-10 -> k=24 "
                       "(test/test_svd.jl:53-83); edge cases er=0 / all<=er",
    "truncate": "relative-er wrapper of truncation_rank; brute force over "
                "all k on small spectra",
    "contract_svd": "closed form t1*t2 at er=0 (test/test_svd.jl:7-11); "
                    "equals einsum on 4-leg tensors (test_svd.jl:13-25); "
                    "error bound of test_svd.jl:75-81",
    "ncon": "einsum definition; brute-force loops on tiny tensors",
    "ncon_ordered": "equals ncon (einsum) under several label orders, with "
                    "traces, parallel legs and an outer product; a chain "
                    "past einsum's 52 letters equals the matrix product",
    "mps_from_state": "round trip MPS(psi) -> contract == psi "
                      "(test/test_mps.jl:34-49)",
    "mpo_from_matrix": "contract(MPO(U)) == U as a (2,..,2) tensor; "
                       "operator-Schmidt closed forms (SWAP, CNOT)",
    "apply_mpo_dense": "dense kron/matvec (test/test_mpo.jl:62-121)",
    "TebdState": "untruncated runs equal the dense statevector simulator; "
                 "product-state closed forms",
    "StateVector": "closed forms: X/H/CNOT on basis states, unitarity",
    "TebdState.switch": "the reference's own case (test/test_mps.jl:198-"
                        "221): switch!(mps, [(2,3),(4,6),(4,5)]) on kron(b..)"
                        " gives kron(b[[1,3,2,5,6,4]]..); wires i, j "
                        "exchange bits n-i and n-j of the dense vector",
    "TebdState.apply_2q_long": "untruncated: equals the dense vector with "
                               "the gate applied on wires (i, j) by "
                               "apply_mpo_dense, both orders of i, j",
    "tfim_mpo": "contracted to a dense matrix it equals the kron-sum TFIM "
                "Hamiltonian (tfim_dense) for n <= 10",
    "expect_mpo": "equals psi^H H psi of the dense vector and matrix for "
                  "n <= 10; <psi|I|psi> = <psi|psi>",
    "canonicalize": "round trip to the dense vector; right-canonical "
                    "sites; lambda = Schmidt values of the dense vector",
    "expect_mpo_blas": "equals expect_mpo (einsum form) on random states",
    "zipup": "without truncation equals the dense H psi (n <= 10); the "
             "truncated result's lambda and norm follow the same rule",
    "mlp_forward": "torch nn.Linear / exact-erf GELU layers built from the "
                   "flat parameter layout, float64 (test_oracle.py)",
    "mlp_loss_grad": "central finite differences of the loss over random "
                     "parameters of every block",
}


# ---------------------------------------------------------------------------
# truncation (src/svd.jl:29-33)
class TruncationError(ValueError):
    """Every singular value lies within the truncation error (the reference
    errors here: findfirst returns nothing)."""


def truncation_rank(s, er):
    """Number of leading singular values kept by contract_svd's rule.

    c = sqrt.(cumsum(reverse(S.^2))); k = length(S) - findfirst(c .> er) + 1
    (src/svd.jl:29-32).  s must be sorted descending.
    """
    s = np.asarray(s, dtype=np.float64)
    c = np.sqrt(np.cumsum(s[::-1] ** 2))
    hit = np.nonzero(c > er)[0]
    if hit.size == 0:
        raise TruncationError("every singular value is within er")
    k_rev = int(hit[0]) + 1
    return s.size - k_rev + 1


def truncate(s, cutoff, relative=True, chimax=0):
    """(k, discarded weight, total weight) for the build's truncation.

    er = cutoff * ||s||_2 when relative (||s||^2 summed from the smallest
    value up, the same order as the tail sums), else er = cutoff; then
    k = min(k, chimax) when chimax > 0.  The discarded weight is the sum of
    the squares of the dropped values, again from the smallest up.
    """
    s = np.asarray(s, dtype=np.float64)
    if cutoff < 0:
        raise ValueError("Error must be positive")
    cs = np.cumsum(s[::-1] ** 2)
    total = float(cs[-1]) if s.size else 0.0
    er = cutoff * np.sqrt(total) if relative else cutoff
    k = truncation_rank(s, er)
    if chimax and chimax > 0:
        k = min(k, int(chimax))
    n = s.size
    disc = float(cs[n - k - 1]) if k < n else 0.0
    return k, disc, total


# ---------------------------------------------------------------------------
# contraction semantics
def ncon(tensors, indexlists):
    """Full contraction with TensorOperations.ncon labelling as produced by
    contract_rep (src/contract.jl:39-60): positive labels are summed,
    negative labels -1, -2, ... give the output axes in that order.
    Evaluated by einsum (a closed form of the definition)."""
    letters = {}
    alphabet = [chr(c) for c in range(ord("a"), ord("z") + 1)] + \
        [chr(c) for c in range(ord("A"), ord("Z") + 1)]

    def let(lbl):
        if lbl not in letters:
            letters[lbl] = alphabet[len(letters)]
        return letters[lbl]

    subs = []
    for lst in indexlists:
        subs.append("".join(let(l) for l in lst))
    outs = sorted({l for lst in indexlists for l in lst if l < 0},
                  reverse=True)
    out = "".join(let(l) for l in outs)
    expr = ",".join(subs) + "->" + out
    return np.einsum(expr, *[np.asarray(t) for t in tensors], optimize=True)


def ncon_ordered(tensors, indexlists, order=None):
    """ncon as TensorOperations evaluates it: the positive labels
    contracted one at a time in `order` (default: increasing), each step a
    tensordot of the two tensors carrying the label over every label they
    share (a trace when one tensor carries both ends); tensors left
    unconnected at the end are joined by outer products in their order;
    output axes -1, -2, ....  For networks with more labels than einsum's
    alphabet."""
    items = [(np.asarray(t), list(l)) for t, l in zip(tensors, indexlists)]
    if order is None:
        order = sorted({x for l in indexlists for x in l if x > 0})
    for lbl in order:
        holders = [i for i, (_, l) in enumerate(items) if lbl in l]
        if not holders:
            continue  # contracted with an earlier label
        if len(holders) == 1:
            t, l = items[holders[0]]
            a, b = [k for k, x in enumerate(l) if x == lbl]
            t = np.trace(t, axis1=a, axis2=b)
            items[holders[0]] = (t, [x for k, x in enumerate(l)
                                     if k not in (a, b)])
            continue
        i, j = holders
        (ta, la), (tb, lb) = items[i], items[j]
        common = [x for x in la if x > 0 and x in lb]
        ax_a = [la.index(x) for x in common]
        ax_b = [lb.index(x) for x in common]
        t = np.tensordot(ta, tb, axes=(ax_a, ax_b))
        l = [x for x in la if x not in common] + \
            [x for x in lb if x not in common]
        items[i] = (t, l)
        del items[j]
    t, l = items[0]
    for t2, l2 in items[1:]:
        t = np.multiply.outer(t, t2)
        l = l + l2
    outs = sorted([x for x in l if x < 0], reverse=True)
    return np.transpose(t, [l.index(x) for x in outs]) if outs else t


def contract_network(tensors, contractions, openidx):
    """contract(net) of src/contract.jl:242-264 (non-optimised order).

    contractions: list of lists of (tensor, leg) pairs, 1-based;
    openidx: list of (tensor, leg), 1-based.  Output axes follow openidx.
    """
    if len(tensors) == 1:
        perm = [leg - 1 for (_, leg) in openidx]
        return np.transpose(np.asarray(tensors[0]), perm)
    index = [[0] * np.ndim(t) for t in tensors]
    for i, summ in enumerate(contractions, start=1):
        for (t, leg) in summ:
            index[t - 1][leg - 1] = i
    J = len(contractions)
    for i, (t, leg) in enumerate(openidx, start=1):
        assert index[t - 1][leg - 1] == 0
        index[t - 1][leg - 1] = -i - J
    for lst in index:
        assert all(l != 0 for l in lst)
    # relabel open legs to -1, -2, ... preserving order
    return ncon(tensors, [[(l + J if l < 0 else l) for l in lst]
                          for lst in index])


def contract_pair(A, B, legsA, legsB, outperm=None):
    """One pairwise ncon step: contract legsA of A with legsB of B (1-based);
    natural output = free legs of A then free legs of B; then
    permutedims(natural, outperm)."""
    la = [l - 1 for l in legsA]
    lb = [l - 1 for l in legsB]
    C = np.tensordot(np.asarray(A), np.asarray(B), axes=(la, lb))
    if outperm is not None:
        C = np.transpose(C, [p - 1 for p in outperm])
    return C


def contract_svd(T1, T2, indx, er=0.0):
    """src/svd.jl:7-38: SVD both operands, truncate each with the absolute
    tail-norm rule, multiply back."""
    if er < 0:
        raise ValueError("Error must be positive")
    i1, i2 = indx
    T1 = np.asarray(T1)
    T2 = np.asarray(T2)
    n1, n2 = T1.ndim, T2.ndim
    # Julia's size(A, d) is 1 for d > ndims(A)
    D1 = T1.shape[i1 - 1] if i1 <= n1 else 1
    D2 = T2.shape[i2 - 1] if i2 <= n2 else 1
    if D1 != D2:
        raise ValueError("Dimensions of contraction legs do not match")
    s1, s2 = T1.shape, T2.shape
    newdim = list(s1[:i1 - 1]) + list(s1[i1:]) + list(s2[:i2 - 1]) + \
        list(s2[i2:])
    p1 = list(range(i1 - 1)) + list(range(i1, n1)) + [i1 - 1]
    p2 = [i2 - 1] + list(range(i2 - 1)) + list(range(i2, n2))
    T1p = np.transpose(T1, p1).reshape((-1, D1), order="F")
    T2p = np.transpose(T2, p2).reshape((D1, -1), order="F")
    U1, S1, V1h = np.linalg.svd(T1p, full_matrices=False)
    U2, S2, V2h = np.linalg.svd(T2p, full_matrices=False)
    k1 = truncation_rank(S1, er)
    k2 = truncation_rank(S2, er)
    T = (U1[:, :k1] * S1[:k1]) @ V1h[:k1] @ (U2[:, :k2] * S2[:k2]) @ V2h[:k2]
    return T.reshape(newdim, order="F")


# ---------------------------------------------------------------------------
# MPS(psi) and MPO(m) (exact SVD sweeps)
def mps_from_state(psi):
    """src/mps.jl:55-89: statevector -> list of site tensors (first (2,chi),
    middle (chi,2,chi'), last (chi,2)), no truncation."""
    psi = np.asarray(psi, dtype=np.complex128)
    n = psi.size
    if n == 0 or (n & (n - 1)) != 0:
        raise ValueError("Input state must have length 2^N")
    M = int(np.log2(n))
    tensors = []
    m = psi.reshape((2, 2 ** (M - 1)), order="F")
    U, S, Vh = np.linalg.svd(m, full_matrices=False)
    tensors.append(U)
    lbond = S.size
    m = S[:, None] * Vh
    for bit in range(2, M):
        m = m.reshape((lbond * 2, 2 ** (M - bit)), order="F")
        U, S, Vh = np.linalg.svd(m, full_matrices=False)
        rbond = S.size
        m = S[:, None] * Vh
        tensors.append(U.reshape((lbond, 2, rbond), order="F"))
        lbond = rbond
    tensors.append(m)
    return tensors


def mps_to_state(tensors):
    """Contract a closed MPS (first (2,chi), last (chi,2)) to the vector."""
    v = np.asarray(tensors[0])  # (2, chi)
    for t in tensors[1:-1]:
        # v: (2^k, chi) ; t: (chi, 2, chi')
        v = np.einsum("xa,asb->xsb", v, t)
        v = v.reshape((-1, t.shape[2]), order="F")
    v = np.einsum("xa,as->xs", v, np.asarray(tensors[-1]))
    return v.reshape(-1, order="F")


def mpo_from_matrix(m):
    """src/mpo.jl:27-90: operator-Schmidt / TT split of a 2^M x 2^M matrix.
    Returns (tensors, singular values of every split).  Tensors are
    (row, col, chi), (chi, row, col, chi), ..., (chi, row, col)."""
    m = np.asarray(m, dtype=np.complex128)
    M = int(round(np.log2(m.shape[0])))
    if M < 1:
        raise ValueError("Need at least one qubit to act on.")
    if M == 1:
        return [m.reshape((2, 2), order="F")], []
    t = m.reshape((2,) * (2 * M), order="F")
    dims = []
    for i in range(M):
        dims += [i, i + M]
    t = np.transpose(t, dims)
    bond = 1
    mm = t.reshape((4 * bond, -1), order="F")
    U, S, Vh = np.linalg.svd(mm, full_matrices=False)
    spectra = [S]
    bond = S.size
    mm = S[:, None] * Vh
    out = [U.reshape((2, 2, bond), order="F")]
    for i in range(2, M):
        mm = mm.reshape((bond * 4, -1), order="F")
        U, S, Vh = np.linalg.svd(mm, full_matrices=False)
        spectra.append(S)
        nb = S.size
        mm = S[:, None] * Vh
        out.append(U.reshape((bond, 2, 2, nb), order="F"))
        bond = nb
    out.append(mm.reshape((bond, 2, 2), order="F"))
    return out, spectra


def operator_schmidt(g):
    """Singular values of the first split of MPO(g) for a 4x4 gate g (the
    operator-Schmidt coefficients, src/decompose.jl:16-29)."""
    _, spectra = mpo_from_matrix(g)
    return spectra[0]


def apply_mpo_dense(psi, m, iwire):
    """Dense apply_MPO(psi, m, iwire) (src/mpo.jl:228-252): the M-qubit
    matrix m acts on wires iwire (1-based, any order); gate-matrix bit k-1
    is wire iwire[k]."""
    psi = np.asarray(psi, dtype=np.complex128)
    n = int(round(np.log2(psi.size)))
    M = len(iwire)
    t = psi.reshape((2,) * n, order="F")  # axis w-1 = wire w
    g = np.asarray(m).reshape((2,) * (2 * M), order="F")
    # g axes: out bits 0..M-1 then in bits 0..M-1
    in_axes = [w - 1 for w in iwire]
    out = np.tensordot(g, t, axes=(list(range(M, 2 * M)), in_axes))
    # out axes: gate out bits (M) then remaining state axes in order
    rest = [a for a in range(n) if a not in in_axes]
    order = [0] * n
    for k, w in enumerate(in_axes):
        order[w] = k
    for j, a in enumerate(rest):
        order[a] = M + j
    out = np.transpose(out, order)
    return out.reshape(-1, order="F")


# SWAP on index s_a + 2 s_b: |01> <-> |10>
SWAP_GATE = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0],
                      [0, 0, 0, 1]], dtype=np.complex128)


# ---------------------------------------------------------------------------
# dense statevector (independent check of the MPS path)
class StateVector:
    def __init__(self, n):
        self.n = n
        self.psi = np.zeros(2 ** n, dtype=np.complex128)
        self.psi[0] = 1.0

    def apply_1q(self, w, u):
        self.psi = apply_mpo_dense(self.psi, u, (w,))

    def apply_2q(self, w, g):
        self.psi = apply_mpo_dense(self.psi, g, (w, w + 1))

    def apply_2q_any(self, i, j, g):
        """g (index s_i + 2 s_j) on wires i != j."""
        self.psi = apply_mpo_dense(self.psi, g, (i, j))

    def amplitudes(self, bits):
        bits = np.asarray(bits, dtype=np.int64)
        idx = (bits * (1 << np.arange(self.n))).sum(axis=1)
        return self.psi[idx]

    def expect_z(self):
        p = np.abs(self.psi) ** 2
        idx = np.arange(p.size)
        return np.array([np.sum(p * (1 - 2 * ((idx >> w) & 1)))
                         for w in range(self.n)]) / p.sum()


# ---------------------------------------------------------------------------
# TEBD in the right-canonical (Hastings) gauge
class TebdState:
    """|0...0> of n qubits; B[j] has shape (chi_{j-1}, 2, chi_j) (0-based
    list index j-1), lam[j] is the spectrum of bond j (lam[0] = lam[n] =
    [1]).  Same algebra as libqtn's qtn_mps_apply_2q."""

    def __init__(self, n):
        self.n = n
        self.B = []
        for _ in range(n):
            t = np.zeros((1, 2, 1), dtype=np.complex128)
            t[0, 0, 0] = 1.0
            self.B.append(t)
        self.lam = [np.ones(1) for _ in range(n + 1)]

    def copy(self):
        c = TebdState.__new__(TebdState)
        c.n = self.n
        c.B = [b.copy() for b in self.B]
        c.lam = [l.copy() for l in self.lam]
        return c

    def bond_dims(self):
        return [1] + [b.shape[2] for b in self.B]

    def apply_1q(self, w, u):
        """u acts on the physical leg of site w (1-based)."""
        self.B[w - 1] = np.einsum("uv,avb->aub", np.asarray(u),
                                  self.B[w - 1])

    def theta(self, i, g):
        """(Phi, theta) matrices for gate g on sites (i, i+1)."""
        Bl, Br = self.B[i - 1], self.B[i]
        chiL, chiM, chiR = Bl.shape[0], Bl.shape[2], Br.shape[2]
        t0 = Bl.reshape((2 * chiL, chiM), order="F") @ \
            Br.reshape((chiM, 2 * chiR), order="F")
        t4 = t0.reshape((chiL, 2, 2, chiR), order="F")
        G = np.asarray(g).reshape((2, 2, 2, 2), order="F")  # [s',t',s,t]
        p4 = np.einsum("uvst,astb->auvb", G, t4)
        phi = p4.reshape((2 * chiL, 2 * chiR), order="F")
        lamL = np.repeat(self.lam[i - 1][None, :], 2, axis=0).reshape(-1)
        theta = lamL[:, None] * phi
        return phi, theta

    def apply_2q(self, i, g, cutoff, chimax, svd=None):
        """Gate g (4x4, index s_i + 2 s_{i+1}) on sites (i, i+1).
        Returns (k, discarded weight / total, full spectrum)."""
        Bl, Br = self.B[i - 1], self.B[i]
        chiL, chiR = Bl.shape[0], Br.shape[2]
        phi, theta = self.theta(i, g)
        if svd is None:
            _, S, Vh = np.linalg.svd(theta, full_matrices=False)
        else:
            _, S, Vh = svd(theta)
        k, disc, total = truncate(S, cutoff, True, chimax)
        kw = float(np.sum(S[:k][::-1] ** 2))
        nrm = np.sqrt(kw)
        Vk = Vh[:k].conj().T
        self.B[i] = Vh[:k].reshape((k, 2, chiR), order="F")
        self.B[i - 1] = ((phi @ Vk) / nrm).reshape((chiL, 2, k), order="F")
        self.lam[i] = S[:k] / nrm
        return k, (disc / total if total > 0 else 0.0), S

    def apply_2q_long(self, i, j, g, cutoff, chimax):
        """g (4x4, index s_i + 2 s_j) on sites i != j by SWAP routing, the
        device's rule (qtn_mps_apply_2q_long): the qubit of the farther site
        travels to next to the nearer one by SWAP updates, the gate acts on
        (lo, lo + 1) with its bits reordered when i > j, and the qubit
        travels back; every update is apply_2q with cutoff / chimax.
        Returns the gate update's (k, disc, spectrum)."""
        lo, hi = min(i, j), max(i, j)
        g = np.asarray(g, dtype=np.complex128)
        if i > j:
            P = SWAP_GATE
            g = P @ g @ P
        for k in range(hi - 1, lo, -1):
            self.apply_2q(k, SWAP_GATE, cutoff, chimax)
        out = self.apply_2q(lo, g, cutoff, chimax)
        for k in range(lo + 1, hi):
            self.apply_2q(k, SWAP_GATE, cutoff, chimax)
        return out

    def switch(self, i, j):
        """switch!(mps, i, j) (src/switch.jl:63-86): wires numbered from the
        most significant (site n - w + 1), adjacent swaps lo..hi-1 then
        hi-2..lo (src/switch.jl:79-85), untruncated."""
        if not (i > 0 and j > 0):
            raise ValueError("Wire indices `i` and `j` must be positive")
        if not (i <= self.n and j <= self.n):
            raise ValueError("Indices to swap `i` and `j` must be less than "
                             "or equal to the number of open wires in MPS")
        if i == j:
            return
        a, b = self.n - i + 1, self.n - j + 1
        lo, hi = min(a, b), max(a, b)
        for k in range(lo, hi):
            self.apply_2q(k, SWAP_GATE, 0.0, 0)
        for k in range(hi - 2, lo - 1, -1):
            self.apply_2q(k, SWAP_GATE, 0.0, 0)

    def amplitudes(self, bits):
        bits = np.asarray(bits, dtype=np.int64)
        nb = bits.shape[0]
        v = np.ones((nb, 1), dtype=np.complex128)
        for j, B in enumerate(self.B):
            s = bits[:, j]
            v = np.einsum("ia,iab->ib", v, B[:, s, :].transpose(1, 0, 2))
        return v[:, 0]

    def _right_envs(self):
        # R_{j-1}[a, d] = sum_{s,b,c} B[a,s,b] R_j[b,c] conj(B[d,s,c])
        R = [None] * (self.n + 1)
        R[self.n] = np.ones((1, 1), dtype=np.complex128)
        for j in range(self.n, 0, -1):
            B = self.B[j - 1]
            cl, cr = B.shape[0], B.shape[2]
            T = (B.reshape((2 * cl, cr), order="F") @ R[j]).reshape(
                (cl, 2 * cr), order="F")
            R[j - 1] = T @ B.reshape((cl, 2 * cr), order="F").conj().T
        return R

    def expect_z(self):
        R = self._right_envs()
        E = np.ones((1, 1), dtype=np.complex128)
        out = np.zeros(self.n)
        for j in range(1, self.n + 1):
            B = self.B[j - 1]
            cl, cr = B.shape[0], B.shape[2]
            # T[a, s, b] = sum_d E[a, d] B[d, s, b]
            T = np.stack([E @ B[:, 0, :], E @ B[:, 1, :]], axis=1)
            Y = T.reshape((2 * cl, cr), order="F") @ R[j]
            Bm = B.reshape((2 * cl, cr), order="F").conj()
            prod = Bm * Y
            num = prod[:cl].sum() - prod[cl:].sum()
            nrm = prod.sum()
            out[j - 1] = (num / nrm).real
            E = Bm.T @ T.reshape((2 * cl, cr), order="F")
        return out

    def overlap(self, other):
        """<self|other>."""
        E = np.ones((1, 1), dtype=np.complex128)
        for A, B in zip(self.B, other.B):
            al, ar = A.shape[0], A.shape[2]
            bl, br = B.shape[0], B.shape[2]
            T = np.concatenate([E @ B[:, 0, :], E @ B[:, 1, :]], axis=0)
            E = A.reshape((2 * al, ar), order="F").conj().T @ T
        return complex(E[0, 0])

    def norm2(self):
        return self.overlap(self).real

    def to_statevector(self):
        v = np.ones((1, 1), dtype=np.complex128)  # (2^k, chi)
        for B in self.B:
            v = np.einsum("xa,asb->xsb", v, B).reshape((-1, B.shape[2]),
                                                      order="F")
        return v[:, 0]


def run_layers(state, layers, cutoff, chimax, record=None):
    """Apply brick-wall layers ([(singles, pairs)], see
    qaintensor.circuits) to a TebdState; optionally record per-gate
    (site, k, disc, spectrum)."""
    for singles, pairs in layers:
        for (w, u) in singles:
            state.apply_1q(w, u)
        for (w, g) in pairs:
            k, disc, S = state.apply_2q(w, g, cutoff, chimax)
            if record is not None:
                record.append((w, k, disc, S))
    return state


# ---------------------------------------------------------------------------
# MPO x MPS (the build's zip-up path; src/mpo.jl:184-252 only wires an MPO
# into a network, the compressed application is new)
#
# MPO site tensors W_j are (D_{j-1}, 2 out, 2 in, D_j), the layout of the
# middle tensors of MPO(m) (src/mpo.jl:66); (W psi)[s'] = sum_s W[s', s]
# psi[s] with wire j <-> bit j-1 of the state index.
def tfim_mpo(n, J, h):
    """H = -J sum_i Z_i Z_{i+1} - sum_i h_i X_i as an MPO of bond dim 3:
    bulk W = [[I, 0, 0], [Z, 0, 0], [-h X, -J Z, I]] (row = left index),
    first site its last row, last site its first column."""
    I2 = np.eye(2, dtype=np.complex128)
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Z = np.diag([1.0, -1.0]).astype(np.complex128)
    h = np.asarray(h, dtype=np.float64)
    out = []
    for j in range(n):
        W = np.zeros((3, 2, 2, 3), dtype=np.complex128)
        W[0, :, :, 0] = I2
        W[1, :, :, 0] = Z
        W[2, :, :, 0] = -h[j] * X
        W[2, :, :, 1] = -J * Z
        W[2, :, :, 2] = I2
        if j == 0:
            W = W[2:3]
        if j == n - 1:
            W = W[:, :, :, 0:1]
        out.append(W)
    return out


def tfim_dense(n, J, h):
    """The same Hamiltonian as a dense 2^n x 2^n matrix from Kronecker
    products (wire j <-> bit j-1, i.e. the last kron factor is wire 1)."""
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Z = np.diag([1.0, -1.0]).astype(np.complex128)

    def op(ops):  # ops: {wire: 2x2}
        m = np.ones((1, 1), dtype=np.complex128)
        for w in range(n, 0, -1):
            m = np.kron(m, ops.get(w, np.eye(2)))
        return m

    H = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    for i in range(1, n):
        H -= J * op({i: Z, i + 1: Z})
    for i in range(1, n + 1):
        H -= h[i - 1] * op({i: X})
    return H


def mpo_dense(W):
    """Contract an MPO to its dense matrix [row, col] (row = out bits,
    wire j <-> bit j-1)."""
    M = W[0][0]  # (2 out, 2 in, D1)
    for j in range(1, len(W)):
        T = np.einsum("rcd,dsue->rscue", M, W[j])
        R, S_, Cc, U, E = T.shape
        # new row index r + R s, new column c + C u (earlier wires fastest)
        M = T.reshape((R * S_, Cc * U, E), order="F")
    return M[:, :, 0]


def mps_dense(Bs):
    """Dense vector of an MPS given as site tensors (chi_{j-1}, 2, chi_j)."""
    v = np.ones((1, 1), dtype=np.complex128)
    for B in Bs:
        v = np.einsum("xa,asb->xsb", v, B).reshape((-1, B.shape[2]),
                                                  order="F")
    return v[:, 0]


def expect_mpo(Bs, W, As=None):
    """<A|W|B> by a left-to-right environment sweep (A = B by default):
    L[b, d, b'] = sum conj(A[a, s', b]) L[a, d0, a'] W[d0, s', s, d]
    B[a', s, b']."""
    if As is None:
        As = Bs
    L = np.ones((1, 1, 1), dtype=np.complex128)
    for A, B, Wj in zip(As, Bs, W):
        T = np.einsum("adx,xsy->adsy", L, B)
        T = np.einsum("adsy,dtse->atey", T, Wj)
        L = np.einsum("atb,atey->bey", A.conj(), T)
    return complex(L[0, 0, 0])


def expect_mpo_blas(Bs, W):
    """expect_mpo with every contraction a BLAS matmul (for chi ~ 1e3):
    L (a, d, a') -> T1 = L B -> (a, b', d, s) x W(d, s, t, e) -> sum over
    (a, t) with conj(B)."""
    L = np.ones((1, 1, 1), dtype=np.complex128)
    for B, Wj in zip(Bs, W):
        cl, _, cr = B.shape
        Dl, Dr = Wj.shape[0], Wj.shape[3]
        T1 = L.reshape((cl * Dl, cl), order="F") @ \
            B.reshape((cl, 2 * cr), order="F")        # (a, d | s, b')
        T1 = T1.reshape((cl, Dl, 2, cr), order="F").transpose(0, 3, 1, 2)
        Wp = Wj.transpose(0, 2, 1, 3).reshape((Dl * 2, 2 * Dr), order="F")
        T2 = T1.reshape((cl * cr, Dl * 2), order="F") @ Wp  # (a, b' | t, e)
        T2 = T2.reshape((cl, cr, 2, Dr), order="F").transpose(0, 2, 3, 1)
        L = B.reshape((2 * cl, cr), order="F").conj().T @ \
            T2.reshape((2 * cl, Dr * cr), order="F")
        L = L.reshape((cr, Dr, cr), order="F")
    return complex(L[0, 0, 0])


def random_mps(n, chimax, seed):
    """Complex Gaussian site tensors with bonds min(chimax, 2^j, 2^(n-j)),
    then left-canonicalised by a QR sweep and normalised."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [min(chimax, 2 ** j, 2 ** (n - j)) for j in range(n + 1)]
    Bs = []
    for j in range(n):
        a, b = dims[j], dims[j + 1]
        t = (rng.standard_normal((a, 2, b)) +
             1j * rng.standard_normal((a, 2, b))) / np.sqrt(2.0)
        Bs.append(t)
    return left_canonical(Bs)


def left_canonical(Bs):
    """QR sweep left to right; the last site carries the norm, which is
    then divided out."""
    Bs = [b.copy() for b in Bs]
    for j in range(len(Bs) - 1):
        a, _, b = Bs[j].shape
        Q, R = np.linalg.qr(Bs[j].reshape((2 * a, b), order="F"))
        k = Q.shape[1]
        Bs[j] = Q.reshape((a, 2, k), order="F")
        nb = Bs[j + 1].shape[2]
        Bs[j + 1] = (R @ Bs[j + 1].reshape((b, 2 * nb), order="F")
                     ).reshape((k, 2, nb), order="F")
    Bs[-1] = Bs[-1] / np.linalg.norm(Bs[-1])
    return Bs


def canonicalize(Bs, cutoff=0.0, chimax=0, left_done=False):
    """Right-canonical form with bond spectra, the layout of libqtn's
    device MPS: an SVD sweep left to right (left-orthonormal sites), then
    one right to left truncating each bond with the build's rule.  Returns
    (B, lam, norm): B right-canonical, lam[j] the normalised Schmidt values
    of bond j (lam[0] = lam[n] = [1]), norm = |psi|."""
    Bs = [b.copy() for b in Bs]
    n = len(Bs)
    if not left_done:
        for j in range(n - 1):
            a, _, b = Bs[j].shape
            U, S, Vh = np.linalg.svd(Bs[j].reshape((2 * a, b), order="F"),
                                     full_matrices=False)
            k = S.size
            Bs[j] = U.reshape((a, 2, k), order="F")
            nb = Bs[j + 1].shape[2]
            Bs[j + 1] = ((S[:, None] * Vh) @
                         Bs[j + 1].reshape((b, 2 * nb), order="F")
                         ).reshape((k, 2, nb), order="F")
    lam = [None] * (n + 1)
    lam[0] = np.ones(1)
    lam[n] = np.ones(1)
    for j in range(n - 1, 0, -1):
        a, _, b = Bs[j].shape
        U, S, Vh = np.linalg.svd(Bs[j].reshape((a, 2 * b), order="F"),
                                 full_matrices=False)
        k, _, _ = truncate(S, cutoff, True, chimax)
        nrm = np.sqrt(np.sum(S[:k][::-1] ** 2))
        Bs[j] = Vh[:k].reshape((k, 2, b), order="F")
        lam[j] = S[:k] / nrm
        pa = Bs[j - 1].shape[0]
        Bs[j - 1] = (Bs[j - 1].reshape((2 * pa, a), order="F") @
                     (U[:, :k] * S[:k])).reshape((pa, 2, k), order="F")
    norm = np.linalg.norm(Bs[0])
    Bs[0] = Bs[0] / norm
    return Bs, lam, float(norm)


def zipup(Bs, W, cutoff, chimax):
    """Zip-up of W|psi> (Stoudenmire & White, New J. Phys. 12, 055026
    (2010)) left to right with the build's truncation at every step, then
    the right-to-left recompression of canonicalize.  Returns
    (B, lam, norm) of the compressed W|psi>.  The truncations are near
    optimal only for a right-canonical psi (the device MPS gauge)."""
    n = len(Bs)
    C = np.ones((1, 1, 1), dtype=np.complex128)  # (k, D, chi)
    out = []
    for j in range(n):
        B, Wj = Bs[j], W[j]
        k, D, chi = C.shape
        cr, Dr = B.shape[2], Wj.shape[3]
        T1 = np.einsum("kdx,xsb->kdsb", C, B)
        T2 = np.einsum("kdsb,dtse->ktbe", T1, Wj)  # (k, s', b, d')
        if j == n - 1:
            out.append(T2.reshape((k, 2, cr * Dr), order="F"))
            break
        M = T2.reshape((2 * k, cr * Dr), order="F")
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        kk, _, _ = truncate(S, cutoff, True, chimax)
        out.append(U[:, :kk].reshape((k, 2, kk), order="F"))
        C = (S[:kk, None] * Vh[:kk]).reshape((kk, cr, Dr), order="F")
        C = np.transpose(C, (0, 2, 1))
    return canonicalize(out, cutoff, chimax, left_done=True)


# ---------------------------------------------------------------------------
# Rank-predictor MLP (the build's learned part; no counterpart in the
# reference).  Parameters are one flat float array, column-major blocks:
#   W1 (width x nin), b1 (width), W2 (width x width), b2 (width),
#   W3 (1 x width), b3 (1).
# Inputs are columns x_b (nin); features 0..63 enter as
# log10(max(x, 1e-16)) / 8.
def gelu(z):
    from scipy.special import erf
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    from scipy.special import erf
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + \
        z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def mlp_unpack(p, nin, width):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + width * nin].reshape((width, nin), order="F"); o += width * nin
    b1 = p[o:o + width]; o += width
    W2 = p[o:o + width * width].reshape((width, width), order="F")
    o += width * width
    b2 = p[o:o + width]; o += width
    W3 = p[o:o + width].reshape((1, width), order="F"); o += width
    b3 = p[o:o + 1]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def mlp_nparams(nin, width):
    return width * nin + width + width * width + width + width + 1


def mlp_prepare(X, nlog=64):
    X = np.array(X, dtype=np.float64)
    X[:nlog] = np.log10(np.maximum(X[:nlog], 1e-16)) / 8.0
    return X


def mlp_forward(p, X, nin, width):
    """Y (1 x B) for inputs X (nin x B), float64 arithmetic."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, nin, width)
    A0 = mlp_prepare(X)
    Z1 = W1 @ A0 + b1[:, None]
    A1 = gelu(Z1)
    Z2 = W2 @ A1 + b2[:, None]
    A2 = gelu(Z2)
    return W3 @ A2 + b3[:, None]


def mlp_loss_grad(p, X, T, nin, width):
    """(loss, flat gradient) of the mean squared error 1/B sum (y - t)^2."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, nin, width)
    B = X.shape[1]
    A0 = mlp_prepare(X)
    Z1 = W1 @ A0 + b1[:, None]
    A1 = gelu(Z1)
    Z2 = W2 @ A1 + b2[:, None]
    A2 = gelu(Z2)
    Y = W3 @ A2 + b3[:, None]
    R = Y - np.asarray(T, dtype=np.float64).reshape(1, -1)
    loss = float(np.sum(R * R) / B)
    dY = 2.0 * R / B
    dW3 = dY @ A2.T
    db3 = dY.sum(axis=1)
    dZ2 = (W3.T @ dY) * gelu_grad(Z2)
    dW2 = dZ2 @ A1.T
    db2 = dZ2.sum(axis=1)
    dZ1 = (W2.T @ dZ2) * gelu_grad(Z1)
    dW1 = dZ1 @ A0.T
    db1 = dZ1.sum(axis=1)
    g = np.concatenate([dW1.reshape(-1, order="F"), db1,
                        dW2.reshape(-1, order="F"), db2,
                        dW3.reshape(-1, order="F"), db3])
    return loss, g
