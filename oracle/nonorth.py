"""Brute-force matrix elements between fragment product states built from
NON-ORTHOGONAL orbitals, to first order in the inter-fragment overlap.
TEST INFRASTRUCTURE ONLY: pins the hermitian-XRCC diagram restatement
(oracle/hermitian_ref.py, hermitian-XRCC/diagrams/S{,T,U,V}_*.py) by what the
mathematics fixes, independently of how the diagrams factor.

Two fragments with d0 and d1 spin orbitals (global index: fragment 0 first,
alpha before beta inside a fragment).  The orbitals overlap as
S(lam) = 1 + lam * S_off (S_off couples the fragments only).  In an
orthonormal reference basis b_k (Jordan-Wigner matrices, the sign
convention of general-XRCC/density_tensors.c:80-113):

    a+_p = sum_k B_kp b+_k,   a~_q = sum_k C_qk b_k,
    B^T B = S,  C = B^-1   ({a~_q, a+_p} = delta_qp).

To first order B = 1 + lam X, C = 1 - lam X with X = S_off / 2.  The
Hamiltonian is written with the biorthogonal operators on both sides,

    H = sum_pq h_pq a~+_p a~_q + sum_pqrs v_pqrs a~+_p a~+_q a~_s a~_r,

(a~+_p = (a~_p)^T; with the plain integrals of the orbitals this is the
exact Hamiltonian: sum a~+ h a~ = sum a+ (S^-1 h S^-1) a), so the operator's
own indices never meet S and every lam-dependence of a matrix element comes
from the overlap of the remaining strings -- the structure of the
hermitian-XRCC diagrams, whose s01 terms are the lam^1 coefficients.  Every
vector is propagated as a dual number (v0 + lam v1), so the lam^0 and lam^1
coefficients of

    <P|Q>,  <P| a~+_p a~_q |Q>,  <P| a~+_p a~+_q a~_s a~_r |Q>

come out exactly.  Product determinants are |P0 P1> = (fragment-1
creators, ascending, applied last) (fragment-0 creators, ascending)|vac>
(inner='f0'), i.e. the bit-string basis state at lam = 0, or with the
fragments' strings in the other order (inner='f1', a sign (-1)^(N0 N1):
fragment 0's string leftmost, the ordering of the hermitian-XRCC product
states).
"""
import itertools

import numpy as np


def _popcount(x):
    return bin(int(x)).count("1")


class Fock:
    def __init__(self, dim):
        self.dim = dim
        n = 1 << dim
        st = np.arange(n)
        self.maps = []
        for p in range(dim):
            has = ((st >> p) & 1) == 1
            sg = np.array([-1.0 if _popcount(s >> (p + 1)) & 1 else 1.0 for s in st])
            self.maps.append((st[has], st[~has], sg))

    def b(self, p, V):
        """b_p on the last axis of V."""
        occ, _, sg = self.maps[p]
        out = np.zeros_like(V)
        out[..., occ ^ (1 << p)] = sg[occ] * V[..., occ]
        return out

    def bdag(self, p, V):
        _, emp, sg = self.maps[p]
        out = np.zeros_like(V)
        tgt = emp | (1 << p)
        out[..., tgt] = sg[tgt] * V[..., emp]
        return out


class Dual:
    """Operators of the non-orthogonal basis on dual vectors (v0, v1)."""

    def __init__(self, dim, S_off):
        self.F = Fock(dim)
        self.dim = dim
        self.X = 0.5 * np.asarray(S_off, dtype=np.float64)

    def create(self, p, v):
        v0, v1 = v
        F, X = self.F, self.X
        d1 = F.bdag(p, v1)
        for k in range(self.dim):
            if X[k, p] != 0.0:
                d1 = d1 + X[k, p] * F.bdag(k, v0)
        return F.bdag(p, v0), d1

    def annihilate(self, q, v):
        """a~_q = sum_k C_qk b_k, C = 1 - lam X."""
        v0, v1 = v
        F, X = self.F, self.X
        d1 = F.b(q, v1)
        for k in range(self.dim):
            if X[q, k] != 0.0:
                d1 = d1 - X[q, k] * F.b(k, v0)
        return F.b(q, v0), d1


def _dot(x, y, axes):
    """Dual contraction over the Fock axis: (x0.y0, x0.y1 + x1.y0)."""
    return (np.tensordot(x[0], y[0], axes), np.tensordot(x[0], y[1], axes) + np.tensordot(x[1], y[0], axes))


def product_dets(d0, d1, N0, N1):
    """(occupations of fragment 0 local, of fragment 1 local) pairs in the
    order (P0 lexical, P1 lexical) -- P0 major."""
    o0 = list(itertools.combinations(range(d0), N0))
    o1 = list(itertools.combinations(range(d1), N1))
    return [(a, b) for a in o0 for b in o1], len(o0), len(o1)


def det_vectors(D, d0, d1, N0, N1, inner="f0"):
    """Dual Fock vectors of the product determinants (rows)."""
    dets, n0, n1 = product_dets(d0, d1, N0, N1)
    n = 1 << D.dim
    out0 = np.zeros((len(dets), n))
    out1 = np.zeros((len(dets), n))
    for i, (p0, p1) in enumerate(dets):
        v = (np.zeros(n), np.zeros(n))
        v[0][0] = 1.0
        g0 = list(p0)
        g1 = [d0 + x for x in p1]
        first, second = (g0, g1) if inner == "f0" else (g1, g0)
        for g in first + second:
            v = D.create(g, v)
        out0[i], out1[i] = v
    return (out0, out1), n0, n1


class Elements:
    """lam^0 and lam^1 coefficients of the overlap, one- and two-body
    matrix elements between the product determinants of the (bra) and (ket)
    electron counts."""

    def __init__(self, d0, d1, S_off, inner="f0"):
        self.d0, self.d1 = d0, d1
        self.D = Dual(d0 + d1, S_off)
        self.inner = inner
        self._vec = {}

    def vecs(self, N0, N1):
        key = (N0, N1)
        if key not in self._vec:
            self._vec[key] = det_vectors(self.D, self.d0, self.d1, N0, N1, self.inner)[0]
        return self._vec[key]

    def overlap(self, bra, ket):
        return _dot(self.vecs(*bra), self.vecs(*ket), ([1], [1]))

    def one_body(self, bra, ket):
        """[P, Q, p, q] = <P| a~+_p a~_q |Q> = (a~_p |P>) . (a~_q |Q>)."""
        d = self.D.dim
        B, K = self.vecs(*bra), self.vecs(*ket)
        ab = [self.D.annihilate(p, B) for p in range(d)]
        ak = [self.D.annihilate(q, K) for q in range(d)]
        A = (np.stack([x[0] for x in ab], 1), np.stack([x[1] for x in ab], 1))     # [P, p, f]
        C = (np.stack([x[0] for x in ak], 1), np.stack([x[1] for x in ak], 1))     # [Q, q, f]
        r0, r1 = _dot(A, C, ([2], [2]))                                          # [P, p, Q, q]
        return r0.transpose(0, 2, 1, 3), r1.transpose(0, 2, 1, 3)

    def two_body(self, bra, ket):
        """[P, Q, p, q, s, r] = <P| a~+_p a~+_q a~_s a~_r |Q>."""
        d = self.D.dim
        B, K = self.vecs(*bra), self.vecs(*ket)
        # bra side: a~_q a~_p |P> for every (p, q)
        bb0 = np.zeros((B[0].shape[0], d, d, B[0].shape[1]))
        bb1 = np.zeros_like(bb0)
        for p in range(d):
            x = self.D.annihilate(p, B)
            for q in range(d):
                y = self.D.annihilate(q, x)
                bb0[:, p, q], bb1[:, p, q] = y
        # ket side: a~_s a~_r |Q> for every (s, r)
        kk0 = np.zeros((K[0].shape[0], d, d, K[0].shape[1]))
        kk1 = np.zeros_like(kk0)
        for r in range(d):
            x = self.D.annihilate(r, K)
            for s in range(d):
                y = self.D.annihilate(s, x)
                kk0[:, s, r], kk1[:, s, r] = y
        r0, r1 = _dot((bb0, bb1), (kk0, kk1), ([3], [3]))                       # [P, p, q, Q, s, r]
        return r0.transpose(0, 3, 1, 2, 4, 5), r1.transpose(0, 3, 1, 2, 4, 5)


def state_transform(M, zb0, zb1, zk0, zk1, n_ket_axes=0):
    """M[P, Q, ...] over product determinants (P = P0 * n1 + P1) ->
    [I0, I1, J0, J1, ...] for fragment states (rows of z over each
    fragment's determinants, lexical)."""
    nb0, nb1 = zb0.shape[1], zb1.shape[1]
    nk0, nk1 = zk0.shape[1], zk1.shape[1]
    rest = M.shape[2:]
    T = M.reshape((nb0, nb1, nk0, nk1) + rest)
    return np.einsum("ap,bq,cr,ds,pqrs...->abcd...", zb0, zb1, zk0, zk1, T, optimize=True)
