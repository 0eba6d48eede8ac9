"""Analytic gradient of the lowest XR eigenvalue (monomer + pair terms, any
number of fragments).

E = lowest eigenvalue of the XR matrix H(z, integrals) (Hellmann-Feynman:
dE = v^T dH v).  The backward pass runs every forward contraction in
reverse on the GPU:

* dE/dH: the reduced matrices of the eigenvector v over the tensor-product
  basis, D_m = sum over the other fragments of v v^T on fragment m (monomer
  blocks) and the charge-diagonal blocks of the same sum on each fragment
  pair (pair blocks, gathered per charge combination by xr_pair_gather);
* every block term is a bilinear form s A W B^T (general-XRCC/
  H_contractions.c), whose adjoint is three FP64 GEMMs giving dE/dA, dE/dB
  and dE/dW (W an integral block);
* density adjoints go back to the CI coefficients through
  xr_density_adjoint / xr_monomer_2e_adjoint (the factorised densities
  run backwards, no atomics).

With lagrangian=True the coefficient gradient is that of the generalised
eigenvalue with the overlap of the tensor-product states, dE = v^T (dH - E dS) v,
which is what a state optimiser needs when the states drift from
orthonormality (the E term of StateSpaceOptimizer/state_gradients.py:230).
The reference computes the bra-side half of it (state_gradients.py:205-243).

XREnergy wraps forward and backward as a torch.autograd.Function whose
inputs are the coefficient tensors and the integral blocks; orbital-rotation
gradients follow by ordinary autograd through the integral transformation
(orbitals.py).
"""
import torch

from ._lib import context, ptr
from .density import conjugate
from .eigen import DENSE_MAX
from .gemm import mm as _mm, permute4
from .xr import precision


def mm(A, Bm, transA=0, transB=0, alpha=1.0, out=None, beta=0.0, split=False, mode=None):
    """2-D GEMM on libxrhip: out = alpha op(A) op(B) + beta out (split-FP32
    when split is set; as the two-electron contractions run in precision
    `mode` when that is given)."""
    return _mm(A.contiguous(), Bm.contiguous(), transA, transB, alpha, out, beta, split, mode)


def bilinear_adjoint(G, A, W, Bm, s, Abar=None, Bbar=None, Wbar=None, split="f64", wbar_beta=1.0,
                     abar_beta=1.0, bbar_beta=1.0):
    """For H = s A W B^T (A: nA x p, W: p x q, B: nB x q) and G = dE/dH
    (nA x nB): Abar += s G B W^T, Bbar += s G^T A W, Wbar += s A^T G B.
    Each triple product is evaluated in the cheaper association; split is
    the precision mode the products run in ('f64', 'split_f32', 'split_i8':
    two-electron terms follow the system's mode).  A beta of 0 overwrites
    its output instead of adding to it (a fresh accumulator needs no zero
    fill)."""
    nA, p = A.shape
    nB, q = Bm.shape
    x = split
    if Abar is not None:
        if nA * nB * q + nA * q * p <= nB * q * p + nA * nB * p:
            mm(mm(G, Bm, mode=x), W, transB=1, alpha=s, out=Abar, beta=abar_beta, mode=x)   # (G B) W^T
        else:
            mm(G, mm(Bm, W, transB=1, mode=x), alpha=s, out=Abar, beta=abar_beta, mode=x)   # G (B W^T)
    if Bbar is not None:
        if nB * nA * p + nB * p * q <= nA * p * q + nB * nA * q:
            mm(mm(G, A, transA=1, mode=x), W, alpha=s, out=Bbar, beta=bbar_beta, mode=x)    # (G^T A) W
        else:
            mm(G, mm(A, W, mode=x), transA=1, alpha=s, out=Bbar, beta=bbar_beta, mode=x)    # G^T (A W)
    if Wbar is not None:
        if p * nA * nB + p * nB * q <= nA * nB * q + p * nA * q:
            mm(mm(A, G, transA=1, mode=x), Bm, alpha=s, out=Wbar, beta=wbar_beta, mode=x)   # (A^T G) B
        else:
            mm(A, mm(G, Bm, mode=x), transA=1, alpha=s, out=Wbar, beta=wbar_beta, mode=x)   # A^T (G B)


class _Acc:
    """Accumulators keyed like their forward tensors.  take() hands out a
    tensor with the beta its writer uses: 0.0 for a new one, left
    uninitialised since its first writer overwrites all of it, then 1.0."""

    def __init__(self, device):
        self.device = device
        self.d = {}

    def take(self, key, shape):
        t = self.d.get(key)
        if t is None:
            t = self.d[key] = torch.empty(shape, dtype=torch.float64, device=self.device)
            return t, 0.0
        return t, 1.0

    def zeroed(self, key, shape):
        """The accumulator for a writer that can only add (a library call
        accumulating in place): zero-filled when new."""
        t, beta = self.take(key, shape)
        if beta == 0.0:
            t.zero_()
        return t


def _flat(R):
    return R.view(R.shape[0] * R.shape[1], -1)


GRAD_TOL = 1e-10


class XRAdjoint:
    """Forward + backward of E(z, integrals) for an XRSystem of any number of
    fragments (monomer and pair terms).  The lowest root comes from the
    matrix-free operator (dense eigendecomposition for small spaces)."""

    def __init__(self, system, group_combos=True):
        if len(system.frags) < 2:
            raise ValueError("the XR gradient needs at least two fragments")
        self.X = system
        # take the charge combinations of a pair that share an integral block
        # together (False: every combination on its own, the same sums in
        # another order)
        self.group_combos = group_combos
        self._diag_vecs = {}

    def forward(self):
        X = self.X
        X.build_densities(keep_sigma=True)
        self.build_count = X.build_count
        self.op = op = X.operator()
        if op.n <= DENSE_MAX:
            M = op.dense()
            w, v = torch.linalg.eigh(0.5 * (M + M.T))
            w, v = w[:1], v[:, :1]
        else:
            # the gradient is first order in the eigenvector error: converge
            # the residual well below the energy tolerance
            w, v = op.lowest(k=1, tol=GRAD_TOL, screen=not X.plain)
        # the root in its charge sector's compact layout (pair densities)
        self.sector = op.found[0] if op.n > DENSE_MAX else None
        self.E = float(w[0])
        self.v = v[:, 0].contiguous()
        return self.E

    def backward(self, lagrangian=True):
        """Returns (zbar, ibar): zbar[m][chg] = dE/dz (n_states x n_configs),
        ibar[key] = dE/d(integral block) for keys ('T', m, n), ('U', a, m, n),
        ('V', m, n, o, p)."""
        if self.X.build_count != self.build_count:
            raise RuntimeError("the XRSystem was rebuilt after this forward; run forward again "
                               "before its backward")
        ctx = context(self.X.device.index)
        # the integral blocks of an XR Hamiltonian are Hermitian, which the
        # monomer sigma adjoint's one-gather path needs
        ctx.call("xr_assume_hermitian_v", 1)
        try:
            with precision(self.X.precision, self.X.device.index):
                return self._backward(lagrangian)
        finally:
            ctx.call("xr_assume_hermitian_v", 0)

    def _backward(self, lagrangian):
        X, E, v, xop = self.X, self.E, self.v, self.op
        dev = X.device
        nf = len(X.frags)
        D = [xop.monomer_density(v, m) for m in range(nf)]
        rbar = [_Acc(dev) for _ in range(nf)]
        ibar = _Acc(dev)
        zbar = [{c: torch.zeros_like(F.blocks[c].coeffs) for c in F.charges} for F in X.frags]
        ctx = context(dev.index)

        def ib(kind, *key):
            """(accumulator of integral block (kind, key), the beta its writer uses)."""
            t = getattr(X.ints, kind)(*key)
            return ibar.take((kind,) + key, t.shape)

        def h_bar(parts, g):
            for kind, key in parts:
                t, beta = ib(kind, *key)
                if beta == 0.0:
                    t.copy_(g)
                else:
                    t.add_(g)

        # ---- monomer blocks: e1 = Rca . h, E2 (sigma), nuclear constant
        for m, F in enumerate(X.frags):
            parts = [("T", (m, m)), ("U", (m, m, m))]
            h = X.ints.h_sum((m, m, "mono"), parts)
            hb = torch.empty_like(h)
            hb_beta = 0.0
            for c in F.charges:
                p = F.pos[c]
                G1 = D[m][p][:, p].contiguous()
                Rca = F.rho["ca"][(c, c)]
                Rb, rb_beta = rbar[m].take(("ca", (c, c)), Rca.shape)
                # e1[a] = sum Rca[a, pq] h[pq]
                mm(G1.view(-1, 1), h.view(1, -1), out=_flat(Rb), beta=rb_beta)
                mm(_flat(Rca), G1.view(-1, 1), transA=1, out=hb.view(-1, 1), beta=hb_beta)
                hb_beta = 1.0
                blk = F.blocks[c]
                Vb = ibar.zeroed(("V", m, m, m, m), X.ints.V(m, m, m, m).shape)
                W = F.E2W.get(c)
                args = (blk.n_orbs, blk.n_core, blk.n_elec, blk.n_states, blk.n_configs, ptr(blk.coeffs),
                        blk.n_states, blk.n_configs, ptr(blk.coeffs), ptr(X.ints.V(m, m, m, m)), ptr(G1),
                        ptr(zbar[m][c]), ptr(zbar[m][c]), ptr(Vb))
                if W is not None:       # the forward's W = Veff Y (same V and coefficients)
                    ctx.call("xr_monomer_2e_adjoint_keep", *args, ptr(W), W.numel())
                else:
                    ctx.call("xr_monomer_2e_adjoint", *args)
            h_bar(parts, hb)

        # ---- pair blocks
        for (m1, m2) in xop.pairs:
            Gb = xop.pair_density(v, m1, m2, sector=self.sector)
            # the pair matrix is built as S + S^T from the 'source' blocks and
            # half of the 'self' ones; Gb is symmetric, so a source block's
            # adjoint counts twice and the transposed combinations none
            # the charge-conserving (d = 0) and the one-electron-transfer
            # d = -1 combinations share their integral blocks: each set is
            # taken together; the others one by one
            same, transfer = [], []
            for k1, k2 in X.pair_charge_combos(m1, m2):
                role = X.combo_role(k1, k2)
                if role == "derived":
                    continue
                f = 1.0 if role == "self" else 2.0
                d = k1[0] - k1[1]
                if self.group_combos and d == 0:
                    same.append((k1, k2, f))
                elif self.group_combos and d == -1:
                    transfer.append((k1, k2, f))
                else:
                    self._pair_adjoint(m1, m2, k1, k2, Gb, rbar, ib, h_bar, f)
            if same:
                self._pair_adjoint_same(m1, m2, same, Gb, rbar, ib)
            if transfer:
                self._pair_adjoint_transfer(m1, m2, transfer, Gb, rbar, ib, h_bar)
            del Gb

        # ---- densities -> coefficients
        for m, F in enumerate(X.frags):
            # a density taken as the conjugate of another passes its adjoint
            # back through the same (self-inverse) conjugation
            for (op, key) in [k for k in rbar[m].d if k in F.derived]:
                cop, ckey = F.derived[(op, key)]
                out, beta = rbar[m].take((cop, ckey), F.rho[cop][ckey].shape)
                conjugate(rbar[m].d.pop((op, key)), out=out, accumulate=beta == 1.0)
            for (op, key), Rb in rbar[m].d.items():
                bb, kb = F.blocks[key[0]], F.blocks[key[1]]
                ctx.call("xr_density_adjoint", op.encode(), kb.n_orbs, kb.n_core, kb.n_elec,
                         bb.n_states, bb.n_configs, ptr(bb.coeffs), kb.n_states, kb.n_configs,
                         ptr(kb.coeffs), ptr(Rb), ptr(zbar[m][key[0]]), ptr(zbar[m][key[1]]))
        if lagrangian:
            for m, F in enumerate(X.frags):
                for c in F.charges:
                    p = F.pos[c]
                    Dc = D[m][p][:, p].contiguous()
                    mm(Dc, F.blocks[c].coeffs, alpha=-2.0 * E, out=zbar[m][c], beta=1.0)
        self.D = D
        return zbar, ibar.d

    def _pair_adjoint(self, m1, m2, k1, k2, Gb, rbar, ib, h_bar, factor=1.0):
        X = self.X
        F1, F2 = X.frags[m1], X.frags[m2]
        (i1c, j1c), (i2c, j2c) = k1, k2
        sp = X.precision
        d = i1c - j1c
        sign2 = float((-1) ** ((F2.n_elec_ref - i2c) % 2))

        def G(swapped):
            g = self.op.pair_gather(Gb, m1, m2, (F1.pos, F2.pos), k1, k2, swapped)
            return g.mul_(factor) if factor != 1.0 else g

        R1 = lambda op: F1.rho[op][k1]
        R2 = lambda op: F2.rho[op][k2]
        def R1b(op):
            t, beta = rbar[m1].take((op, k1), F1.rho[op][k1].shape)
            return _flat(t), beta

        def R2b(op):
            t, beta = rbar[m2].take((op, k2), F2.rho[op][k2].shape)
            return _flat(t), beta

        def adj(G_, A, W_, B_, s_, ab, bb, Wb_, **kw):
            """bilinear_adjoint into the accumulators ab, bb ((tensor, beta) pairs)."""
            bilinear_adjoint(G_, A, W_, B_, s_, ab[0], bb[0], Wb_, abar_beta=ab[1], bbar_beta=bb[1], **kw)
        n1, n2 = F1.dim, F2.dim

        def vperm(key, perm, shape2):
            """W = V[key] with axes permuted and flattened, plus its accumulator
            mapped back through the same permutation (library passes)."""
            Vt = X.ints.V(*key).contiguous()
            W = permute4(Vt, perm).view(*shape2)
            Wb = torch.empty_like(W)          # written whole by its one product (wbar_beta = 0)

            def fin():
                out, beta = ib("V", *key)
                permute4(Wb.view([Vt.shape[i] for i in perm]), _inv(perm), out=out, accumulate=beta == 1.0)
            return W, Wb, fin

        if d in (-2, +2):
            if d == -2:
                W, Wb, fin = vperm((m1, m1, m2, m2), (0, 1, 3, 2), (n1 * n1, n2 * n2))
                adj(G(False), _flat(R1("cc")), W, _flat(R2("aa")), 1.0, R1b("cc"), R2b("aa"), Wb, split=sp,
                    wbar_beta=0.0)
            else:
                W, Wb, fin = vperm((m2, m2, m1, m1), (0, 1, 3, 2), (n2 * n2, n1 * n1))
                adj(G(True), _flat(R2("cc")), W, _flat(R1("aa")), 1.0, R2b("cc"), R1b("aa"), Wb, split=sp,
                    wbar_beta=0.0)
            fin()
            return
        if d in (-1, +1):
            if d == -1:
                # creating fragment 1, annihilating fragment 2
                cre, ann, nc, na, s = (R1, R1b, n1), (R2, R2b, n2), n1, n2, sign2
                parts = [("T", (m1, m2)), ("U", (m1, m1, m2)), ("U", (m2, m1, m2))]
                hkey = (m1, m2, "1min")
                k11, k22 = (m1, m1, m1, m2), (m1, m2, m2, m2)
                Gm = G(False)
            else:
                cre, ann, nc, na, s = (R2, R2b, n2), (R1, R1b, n1), n2, n1, -sign2
                parts = [("T", (m2, m1)), ("U", (m1, m2, m1)), ("U", (m2, m2, m1))]
                hkey = (m2, m1, "1pls")
                k11, k22 = (m2, m2, m2, m1), (m2, m1, m1, m1)
                Gm = G(True)
            Rc, Rcb = cre[0], cre[1]
            Ra, Rab = ann[0], ann[1]
            h = X.ints.h_sum(hkey, parts)
            hb = torch.empty_like(h)
            adj(Gm, _flat(Rc("c")), h, _flat(Ra("a")), s, Rcb("c"), Rab("a"), hb, wbar_beta=0.0)
            h_bar(parts, hb)
            W1, W1b, fin1 = vperm(k11, (1, 0, 2, 3), (nc ** 3, na))
            adj(Gm, _flat(Rc("cca")), W1, _flat(Ra("a")), 2.0 * s, Rcb("cca"), Rab("a"), W1b, split=sp,
                wbar_beta=0.0)
            fin1()
            W2, W2b, fin2 = vperm(k22, (0, 1, 3, 2), (nc, na ** 3))
            adj(Gm, _flat(Rc("c")), W2, _flat(Ra("caa")), 2.0 * s, Rcb("c"), Rab("caa"), W2b, split=sp,
                wbar_beta=0.0)
            fin2()
            return
        # d == 0
        Gm = G(False)
        W, Wb, fin = vperm((m1, m2, m1, m2), (0, 2, 1, 3), (n1 * n1, n2 * n2))
        adj(Gm, _flat(R1("ca")), W, _flat(R2("ca")), 4.0, R1b("ca"), R2b("ca"), Wb, split=sp, wbar_beta=0.0)
        fin()
        self._ext_pot(m1, m2, [(k1, k2, Gm)], rbar, ib)

    def _diag_vec(self, nI, nJ):
        """(nI nJ x 1) indicator of the I == J entries of a state-pair index
        (kept: the same few shapes recur in every backward)."""
        cache = self._diag_vecs
        t = cache.get((nI, nJ))
        if t is None:
            t = torch.zeros(nI * nJ, 1, dtype=torch.float64, device=self.X.device)
            t.view(nI, nJ).diagonal().fill_(1.0)
            cache[(nI, nJ)] = t
        return t

    def _ext_pot(self, m1, m2, items, rbar, ib):
        """extPot terms of charge-conserving combinations [(k1, k2, G)]
        (build_H.py:85-86): e2[b] from the I1 == J1 rows of each pair density
        block, e1[a] from its I2 == J2 columns.  Both enter linearly, so they
        are summed over the combinations first: e2 per k2, e1 per k1, then one
        outer product with U and one product with the density per charge."""
        X = self.X
        F1, F2 = X.frags[m1], X.frags[m2]
        E2, E1 = {}, {}
        for k1, k2, Gm in items:
            d1 = self._diag_vec(F1.n_states[k1[0]], F1.n_states[k1[1]])
            d2 = self._diag_vec(F2.n_states[k2[0]], F2.n_states[k2[1]])
            E2[k2] = mm(Gm, d1, transA=1, out=E2.get(k2), beta=1.0 if k2 in E2 else 0.0)   # (nB x 1)
            E1[k1] = mm(Gm, d2, out=E1.get(k1), beta=1.0 if k1 in E1 else 0.0)             # (nA x 1)
        U12 = X.ints.U(m1, m2, m2)
        U21 = X.ints.U(m2, m1, m1)
        for E, U_, F, m, key in ((E2, U12, F2, m2, ("U", m1, m2, m2)), (E1, U21, F1, m1, ("U", m2, m1, m1))):
            for k, eb in E.items():
                R = F.rho["ca"][k]
                out, beta = rbar[m].take(("ca", k), R.shape)
                mm(eb, U_.view(1, -1), out=_flat(out), beta=beta)
                out, beta = ib(*key)
                mm(_flat(R), eb, transA=1, out=out.view(-1, 1), beta=beta)

    def _vperm(self, key, perm, shape2, ib):
        """W = V[key] with axes permuted and flattened, an uninitialised
        accumulator Wb of its shape, and fin() mapping Wb back through the
        same permutation into the accumulator of V[key] (library passes)."""
        Vt = self.X.ints.V(*key).contiguous()
        W = permute4(Vt, perm).view(*shape2)
        Wb = torch.empty_like(W)

        def fin():
            out, beta = ib("V", *key)
            permute4(Wb.view([Vt.shape[i] for i in perm]), _inv(perm), out=out, accumulate=beta == 1.0)
        return W, Wb, fin

    @staticmethod
    def _group_adjoint(items, A, W, B, s, a_take, b_take, Wb, cache, akey, bkey, mode):
        """bilinear_adjoint summed over charge combinations that share W.
        items: [(k1, k2, g)] with g = dE/dH_c (its combination's factor and
        sign included) for H_c = s A[k1] W B[k2]^T.  Adds
            Abar[k1] += s sum_k2 g B[k2] W^T,   Bbar[k2] += s sum_k1 g^T A[k1] W
        into a_take(k1) / b_take(k2) ((tensor, beta) of the accumulator) and
        writes Wb = s sum A[k1]^T g B[k2].  Each output takes the association
        with fewer flops; the partner sums T[k1] = sum_k2 g B[k2] and
        U[k2] = sum_k1 g^T A[k1] are kept in cache (keyed by the operand's name
        akey / bkey and the precision mode, an FP64 sum serving any mode) for
        the other terms of the same combinations.  Either association never
        needs more flops than taking the combinations one by one."""
        k1s = list(dict.fromkeys(k1 for k1, _, _ in items))
        k2s = list(dict.fromkeys(k2 for _, k2, _ in items))
        p, q = W.shape
        nA = {k: A[k].shape[0] for k in k1s}
        nB = {k: B[k].shape[0] for k in k2s}

        def have(kind, name, ks):
            return all((kind, name, k, "f64") in cache or (kind, name, k, mode) in cache for k in ks)

        def part(kind, name, k):
            """T[k] (kind 'T', over the items with k1 = k) or U[k] (kind 'U')."""
            for m in ("f64", mode):
                if (kind, name, k, m) in cache:
                    return cache[(kind, name, k, m)]
            t = None
            for a, b, g in items:
                if kind == "T" and a == k:
                    t = mm(g, B[b], out=t, beta=0.0 if t is None else 1.0, mode=mode)
                elif kind == "U" and b == k:
                    t = mm(g, A[a], transA=1, out=t, beta=0.0 if t is None else 1.0, mode=mode)
            cache[(kind, name, k, mode)] = t
            return t

        cT = 0 if have("T", bkey, k1s) else sum(nA[a] * nB[b] * q for a, b, _ in items)
        cU = 0 if have("U", akey, k2s) else sum(nA[a] * nB[b] * p for a, b, _ in items)
        # Abar: s T[k1] W^T, or s sum_k2 g (B[k2] W^T)
        if cT + sum(nA[k] * q * p for k in k1s) <= (sum(nB[k] * q * p for k in k2s)
                                                     + sum(nA[a] * nB[b] * p for a, b, _ in items)):
            for k1 in k1s:
                t, beta = a_take(k1)
                mm(part("T", bkey, k1), W, transB=1, alpha=s, out=t, beta=beta, mode=mode)
            cT = 0
        else:
            BW = {k2: mm(B[k2], W, transB=1, mode=mode) for k2 in k2s}
            for k1 in k1s:
                t, beta = a_take(k1)
                for a, b, g in items:
                    if a == k1:
                        mm(g, BW[b], alpha=s, out=t, beta=beta, mode=mode)
                        beta = 1.0
        # Bbar: s U[k2] W, or s sum_k1 g^T (A[k1] W)
        if cU + sum(nB[k] * p * q for k in k2s) <= (sum(nA[k] * p * q for k in k1s)
                                                     + sum(nA[a] * nB[b] * q for a, b, _ in items)):
            for k2 in k2s:
                t, beta = b_take(k2)
                mm(part("U", akey, k2), W, alpha=s, out=t, beta=beta, mode=mode)
            cU = 0
        else:
            AW = {k1: mm(A[k1], W, mode=mode) for k1 in k1s}
            for k2 in k2s:
                t, beta = b_take(k2)
                for a, b, g in items:
                    if b == k2:
                        mm(g, AW[a], transA=1, alpha=s, out=t, beta=beta, mode=mode)
                        beta = 1.0
        # Wbar: s sum_k1 A[k1]^T T[k1], or s sum_k2 U[k2]^T B[k2]
        if cT + sum(p * nA[k] * q for k in k1s) <= cU + sum(p * nB[k] * q for k in k2s):
            for i, k1 in enumerate(k1s):
                mm(A[k1], part("T", bkey, k1), transA=1, alpha=s, out=Wb, beta=0.0 if i == 0 else 1.0, mode=mode)
        else:
            for i, k2 in enumerate(k2s):
                mm(part("U", akey, k2), B[k2], transA=1, alpha=s, out=Wb, beta=0.0 if i == 0 else 1.0, mode=mode)

    def _gathered(self, m1, m2, combos, Gb, sign=None):
        """[(k1, k2, g)]: each combination's block of the pair density times
        its factor (and sign(k2) when given)."""
        F1, F2 = self.X.frags[m1], self.X.frags[m2]
        items = []
        for k1, k2, f in combos:
            g = self.op.pair_gather(Gb, m1, m2, (F1.pos, F2.pos), k1, k2, False)
            c = f * (sign(k2) if sign is not None else 1.0)
            if c != 1.0:
                g.mul_(c)
            items.append((k1, k2, g))
        return items

    def _takers(self, m, F, rbar):
        """op -> (k -> (flattened accumulator of F.rho[op][k], beta))."""
        def for_op(op):
            def take(k):
                t, beta = rbar[m].take((op, k), F.rho[op][k].shape)
                return _flat(t), beta
            return take
        return for_op

    def _pair_adjoint_same(self, m1, m2, combos, Gb, rbar, ib):
        """The charge-conserving combinations (k1 = (a, a), k2 = (b, b)) of
        one pair together: H_c = 4 Rca1[k1] W Rca2[k2]^T with one permuted
        V(m1, m2, m1, m2) for all of them (the d == 0 case of _pair_adjoint),
        plus each combination's extPot terms."""
        X = self.X
        F1, F2 = X.frags[m1], X.frags[m2]
        n1, n2 = F1.dim, F2.dim
        items = self._gathered(m1, m2, combos, Gb)
        W, Wb, fin = self._vperm((m1, m2, m1, m2), (0, 2, 1, 3), (n1 * n1, n2 * n2), ib)
        A = {k1: _flat(F1.rho["ca"][k1]) for k1, _, _ in items}
        B = {k2: _flat(F2.rho["ca"][k2]) for _, k2, _ in items}
        self._group_adjoint(items, A, W, B, 4.0, self._takers(m1, F1, rbar)("ca"),
                            self._takers(m2, F2, rbar)("ca"), Wb, {}, "ca", "ca", X.precision)
        fin()
        self._ext_pot(m1, m2, items, rbar, ib)

    def _pair_adjoint_transfer(self, m1, m2, combos, Gb, rbar, ib, h_bar):
        """The d = -1 combinations (fragment 1 gains the electron fragment 2
        loses) of one pair together: the one-electron term s c1 h a2^T and
        the two-electron terms 2 s cca1 W1 a2^T and 2 s c1 W2 caa2^T (the
        d == -1 case of _pair_adjoint), with s the sign of k2 folded into
        each combination's block.  h, W1 and W2 are shared by all of them,
        and so are the partner sums over a2 and c1."""
        X = self.X
        F1, F2 = X.frags[m1], X.frags[m2]
        n1, n2 = F1.dim, F2.dim
        sp = X.precision
        items = self._gathered(m1, m2, combos, Gb,
                               sign=lambda k2: float((-1) ** ((F2.n_elec_ref - k2[0]) % 2)))
        R1 = lambda op: {k1: _flat(F1.rho[op][k1]) for k1, _, _ in items}
        R2 = lambda op: {k2: _flat(F2.rho[op][k2]) for _, k2, _ in items}
        t1, t2 = self._takers(m1, F1, rbar), self._takers(m2, F2, rbar)
        cache = {}
        parts = [("T", (m1, m2)), ("U", (m1, m1, m2)), ("U", (m2, m1, m2))]
        h = X.ints.h_sum((m1, m2, "1min"), parts)
        hb = torch.empty_like(h)
        self._group_adjoint(items, R1("c"), h, R2("a"), 1.0, t1("c"), t2("a"), hb, cache, "c", "a", "f64")
        h_bar(parts, hb)
        W1, W1b, fin1 = self._vperm((m1, m1, m1, m2), (1, 0, 2, 3), (n1 ** 3, n2), ib)
        self._group_adjoint(items, R1("cca"), W1, R2("a"), 2.0, t1("cca"), t2("a"), W1b, cache, "cca", "a", sp)
        fin1()
        W2, W2b, fin2 = self._vperm((m1, m2, m2, m2), (0, 1, 3, 2), (n1, n2 ** 3), ib)
        self._group_adjoint(items, R1("c"), W2, R2("caa"), 2.0, t1("c"), t2("caa"), W2b, cache, "c", "caa", sp)
        fin2()


def _inv(perm):
    inv = [0] * len(perm)
    for i, p in enumerate(perm):
        inv[p] = i
    return inv


class XREnergy(torch.autograd.Function):
    """E = lowest XR eigenvalue as a function of the coefficient
    tensors and the integral blocks; backward is the hand-written adjoint.

    apply(system, lagrangian, z_keys, int_keys, *tensors): tensors are the
    coefficient tensors for z_keys [(m, chg)] followed by the integral blocks
    for int_keys (IntegralCache keys)."""

    @staticmethod
    def forward(ctx, system, lagrangian, z_keys, int_keys, *tensors):
        nz = len(z_keys)
        for (m, c), t in zip(z_keys, tensors[:nz]):
            F = system.frags[m]
            F.blocks[c] = F.blocks[c].with_coeffs(t.detach())
        for key, t in zip(int_keys, tensors[nz:]):
            system.ints.set_block(key, t)
        adj = XRAdjoint(system)
        E = adj.forward()
        ctx.adj, ctx.lag, ctx.z_keys, ctx.int_keys = adj, lagrangian, z_keys, int_keys
        return torch.tensor(E, dtype=torch.float64, device=system.device)

    @staticmethod
    def backward(ctx, gE):
        zbar, ibar = ctx.adj.backward(lagrangian=ctx.lag)
        g = float(gE)
        scale = (lambda t: t) if g == 1.0 else (lambda t: t * g)    # dE/dE = 1 on the gradient path
        out = [scale(zbar[m][c]) for (m, c) in ctx.z_keys]
        for key in ctx.int_keys:
            t = ibar.get(key)
            out.append(None if t is None else scale(t))
        return (None, None, None, None) + tuple(out)


def energy_and_gradient(system, lagrangian=True):
    """(E, v, zbar, ibar) for an XRSystem without autograd plumbing."""
    adj = XRAdjoint(system)
    E = adj.forward()
    zbar, ibar = adj.backward(lagrangian=lagrangian)
    return E, adj.v, zbar, ibar


DimerAdjoint = XRAdjoint


def herm_state_gradient(zbar, system):
    """The state gradient in the definition of the reference's "herm" level
    (StateSpaceOptimizer/state_gradients.py:121-244, xr_order 0):
    g[m][c][k, p] = sum_l d[k, l] <p l| H - E |Psi>, half the derivative of
    the variational XR energy (every product state with its overlaps), for
    fragment m's states k and determinants p of charge c.

    The adjoint's dE/dz differentiates the XR matrix as assembled (the
    identity on the spectator fragments, general-XRCC/hamiltonian.py:21-83).
    For orthonormal states the two agree on every direction orthogonal to
    each charge block's state span, and the variational energy is
    invariant within the span, so g = (1/2) zbar (1 - z^T z) for zbar
    from either lagrangian setting (the overlap term -2 E D z lies in the
    span).  Returns {(m, c): g} as device tensors."""
    out = {}
    for m, F in enumerate(system.frags):
        for c in F.charges:
            z = F.blocks[c].coeffs
            zb = zbar[m][c]
            out[(m, c)] = 0.5 * (zb - mm(mm(zb, z, transB=1), z))
    return out
