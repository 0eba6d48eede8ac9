"""Batched XR Hamiltonian on the GPU.

XRSystem builds, for a supersystem of fragments, every density the XR blocks
need (one HIP call per operator string and charge pair), the monomer blocks
H1[m], the pair blocks H2[m][n] and optionally the trimer blocks H3, all in
the layouts of general-XRCC/test_H.py:63-108, and assembles the tensor-
product XR matrix (general-XRCC/hamiltonian.py:21-83) on the device.
"""
import ctypes
import math
from contextlib import contextmanager

import numpy as np
import torch

from . import blocks as B
from ._lib import PAIR_1MIN1PLS, PAIR_2MIN2PLS, PAIR_EXEX, PairJob, context, ptr
from .density import OPS_BY_CHG_DIFF, ChargeBlock, build_density_arrays, monomer_2e


PRECISIONS = {"f64": 0, "split_f32": 1, "split_i8": 2}


@contextmanager
def precision(mode, device=0):
    """Runs the enclosed two-electron contractions in `mode` ('f64',
    'split_f32' or 'split_i8') on the device's context, restoring the
    previous mode."""
    ctx = context(device)
    old = ctx.precision
    ctx.set_precision(PRECISIONS[mode] if isinstance(mode, str) else int(mode))
    try:
        yield
    finally:
        ctx.set_precision(old)


class XRFragment:
    """One fragment on the device: its charge blocks, densities and the
    position of every (charge, index) state in the fragment's state list."""

    def __init__(self, frag, device):
        self.device = device
        self.n_orbs, self.n_core, self.n_elec_ref = frag.n_orbs, frag.n_core, frag.n_elec_ref
        self.dim = 2 * frag.n_orbs
        self.charges = list(frag.states)
        self.blocks = {chg: ChargeBlock.from_states(st, frag.n_orbs, frag.n_core, device)
                       for chg, st in frag.states.items()}
        self.state_indices = list(frag.state_indices)
        self.d = len(self.state_indices)
        where = {si: k for k, si in enumerate(self.state_indices)}
        self.pos = {}
        for chg, b in self.blocks.items():
            idx = [where.get((chg, i), -1) for i in range(b.n_states)]
            if min(idx) < 0:
                raise ValueError(f"state_indices misses states of charge {chg}")
            self.pos[chg] = torch.tensor(idx, dtype=torch.int64, device=device)
        self.n_states = {chg: b.n_states for chg, b in self.blocks.items()}
        self.rho = None
        self.derived = {}
        self.E2 = None
        self.E2W = {}

    def build(self, V_mm, keep_sigma=False, conjugates=True, reuse=False):
        """Densities for all charge pairs (not the 4-index one) and the
        monomer two-electron matrices; keep_sigma keeps their intermediate
        W = Veff Y per charge (E2W) for the gradient step; conjugates=False
        builds the lowering strings directly instead of as conjugates;
        reuse writes the densities into the tensors of the previous build
        (fixed addresses for a captured build).  keep_sigma may be a byte
        budget per charge block: an intermediate larger than that is not
        stored (the adjoint recomputes it)."""
        self.derived = {}
        self.rho = build_density_arrays(self.blocks, OPS_BY_CHG_DIFF, skip=("ccaa",), derived=self.derived,
                                        conjugates=conjugates, out=self.rho if reuse else None)
        self.E2, self.E2W = {}, {}
        for chg, b in self.blocks.items():
            if keep_sigma and keep_sigma is not True and self.sigma_bytes(chg) > keep_sigma:
                self.E2[chg] = monomer_2e(b, b, V_mm)        # recomputed by the adjoint instead
            elif keep_sigma:
                self.E2[chg], self.E2W[chg] = monomer_2e(b, b, V_mm, keep=True)
            else:
                self.E2[chg] = monomer_2e(b, b, V_mm)

    def sigma_bytes(self, chg):
        """Bytes of the monomer sigma intermediate W = Veff Y of one charge
        block (kept for the gradient step when the budget allows)."""
        b = self.blocks[chg]
        if b.n_elec < 2:
            return 0
        return 8 * math.comb(b.dim, 2) * math.comb(b.dim, b.n_elec - 2) * b.n_states

    def parity(self):
        return torch.tensor([(self.n_elec_ref - chg) % 2 for chg, _ in self.state_indices],
                            dtype=torch.int8, device=self.device)


def sector_layout(ch1, ch2, segq):
    """Row of pair (a, b) inside its fixed-charge block q = ch1[a] + ch2[b],
    for blocks ordered by increasing a * len(ch2) + b: off[q][a] + rank2[b],
    with rank2[b] the rank of b among the states of its charge and off[q][a]
    the number of charge-q pairs whose first state comes before a."""
    ch1, ch2 = np.asarray(ch1), np.asarray(ch2)
    rank2 = np.zeros(ch2.size, dtype=np.int64)
    n2 = {}
    for c in np.unique(ch2):
        idx = np.nonzero(ch2 == c)[0]
        rank2[idx] = np.arange(idx.size)
        n2[int(c)] = idx.size
    off = {}
    for q in segq:
        sizes = np.array([n2.get(int(q) - int(c), 0) for c in ch1], dtype=np.int64)
        off[q] = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return off, rank2


class XRSystem:
    """Device-resident XR model of a supersystem (fragments with .n_orbs,
    .n_core, .n_elec_ref, .states, .state_indices; fragment-blocked
    integrals with T, U, V; nuclear repulsion matrix)."""

    def __init__(self, fragments, integrals, nuc, device=0, precision="f64", plain=False, sigma_budget=None):
        """precision: 'f64', or 'split_f32' / 'split_i8' for the
        two-electron contraction GEMMs (hi/lo FP32 matrix-core products
        accumulated in FP64, or int8 digit-plane products accumulated
        exactly in int32).

        plain=True is the unbatched FP64 route kept as a comparison point:
        every density string built directly (no conjugates), the monomer
        sigma recomputed in the adjoint, dense pair matrices cut into the
        operator's blocks, one GEMM launch per pair product of the operator
        (no grouped GEMM) and every charge sector converged (no screening).

        sigma_budget: bytes up to which a charge block's monomer sigma
        intermediate is stored for the gradient step (default a quarter of
        the device's memory); a larger one is recomputed in the adjoint."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.plain = bool(plain)
        self.device = torch.device("cuda", device)
        self.frags = [XRFragment(f, self.device) for f in fragments]
        self.ints = B.IntegralCache(integrals, self.device)
        self.nuc = np.asarray(nuc, dtype=np.float64)
        self.built = False
        self.build_count = 0     # densities builds so far (a backward checks it matches its forward)
        self._sector_idx = {}    # per pair: fixed-charge block layout and scatter indices
        if sigma_budget is None:
            sigma_budget = torch.cuda.get_device_properties(self.device).total_memory // 4
        self.sigma_budget = int(sigma_budget)

    def build_densities(self, keep_sigma=False, reuse=False):
        with precision(self.precision, self.device.index):
            for m, F in enumerate(self.frags):
                keep = self.sigma_budget if (keep_sigma and not self.plain) else 0
                F.build(self.ints.V(m, m, m, m), keep, conjugates=not self.plain, reuse=reuse)
        self.built = True
        self.build_count += 1

    def _need(self):
        if not self.built:
            self.build_densities()

    # -- blocks -------------------------------------------------------------
    def monomer_matrix(self, m):
        """H1[m][I,J] = build_matrix_elements.monomer (build_H.py:42-55)."""
        self._need()
        F = self.frags[m]
        out = torch.zeros(F.d, F.d, dtype=torch.float64, device=self.device)
        h = self.ints.h_sum((m, m, "mono"), [("T", (m, m)), ("U", (m, m, m))])
        for chg in F.charges:
            n = F.n_states[chg]
            e1 = B.monomer_1e(F.rho["ca"][(chg, chg)], h)
            p = F.pos[chg]
            B.scatter(out, F.d, [(n, n)], [p], [p], [F.d], e1, [1])
            B.scatter(out, F.d, [(n, n)], [p], [p], [F.d], F.E2[chg], [1])
            if self.nuc[m, m] != 0.0:
                B.scatter(out, F.d, [(n, n)], [p], [p], [F.d], None, [0], diag=[1], scale=self.nuc[m, m])
        return out

    def pair_charge_combos(self, m1, m2):
        F1, F2 = self.frags[m1], self.frags[m2]
        for c1b in F1.charges:
            for c1k in F1.charges:
                if abs(c1b - c1k) > 2:
                    continue
                for c2b in F2.charges:
                    c2k = c1b + c2b - c1k
                    if c2k in F2.charges:
                        yield (c1b, c1k), (c2b, c2k)

    @staticmethod
    def combo_role(k1, k2):
        """How the pair block of charges (k1, k2) is built: 'self' when it is
        its own transpose (both fragments keep their charges; built at half
        weight and completed by A + A^T), 'source' when it is built and its
        transposed combination ((k1[1], k1[0]), (k2[1], k2[0])) is not,
        'derived' for that transposed one (the XR pair matrix is symmetric
        for real states)."""
        kt = ((k1[1], k1[0]), (k2[1], k2[0]))
        if kt == (k1, k2):
            return "self"
        return "source" if (k1, k2) < kt else "derived"

    def _pair_terms(self, m1, m2, place):
        """Adds the pair terms of build_matrix_elements.dimer (build_H.py:56-102)
        into the matrices chosen by place(k1, k2) -> (out, ldo, rows, cols, d)
        (the arguments of blocks.scatter for that charge combination).  Only
        the 'self' (at half weight) and 'source' combinations are built: the
        caller completes each matrix as A + A^T."""
        self._need()
        F1, F2 = self.frags[m1], self.frags[m2]
        for k1, k2 in self.pair_charge_combos(m1, m2):
            role = self.combo_role(k1, k2)
            if role == "derived":
                continue
            half = 0.5 if role == "self" else 1.0
            out, ldo, rows, cols, dims = place(k1, k2)
            nI1, nJ1 = F1.n_states[k1[0]], F1.n_states[k1[1]]
            nI2, nJ2 = F2.n_states[k2[0]], F2.n_states[k2[1]]
            fr = [(nI1, nJ1), (nI2, nJ2)]
            terms, extras = B.dimer_block(m1, m2, F2.n_elec_ref, F1.rho, F2.rho, k1, k2, self.ints)
            nA, nB = nI1 * nJ1, nI2 * nJ2
            for Hb, swapped, scale in terms:
                strides = [1, nA] if swapped else [nB, 1]
                B.scatter(out, ldo, fr, rows, cols, dims, Hb, strides, scale=half * scale)
            for kind, v in extras:
                if kind == "e2":   # where I1 == J1
                    B.scatter(out, ldo, fr, rows, cols, dims, v, [0, 1], diag=[1, 0], scale=half)
                else:              # "e1", where I2 == J2
                    B.scatter(out, ldo, fr, rows, cols, dims, v, [1, 0], diag=[0, 1], scale=half)
            if k1[0] == k1[1] and k2[0] == k2[1] and self.nuc[m1, m2] != 0.0:
                B.scatter(out, ldo, fr, rows, cols, dims, None, [0, 0], diag=[1, 1], scale=half * self.nuc[m1, m2])

    def _pair_job(self, m1, m2, k1, k2, half, out, ldo, rows, cols, mul, keep):
        """The xr_pair_job of one charge combination: the family and operands
        build_matrix_elements.dimer (build_H.py:56-102) uses for it, as
        blocks.dimer_block selects them."""
        F1, F2 = self.frags[m1], self.frags[m2]
        I = self.ints
        d = k1[0] - k1[1]
        sign2 = float((-1) ** ((F2.n_elec_ref - k2[0]) % 2))

        def P(t):
            if not t.is_contiguous():
                raise ValueError("pair job operands must be contiguous")
            keep.append(t)
            return t.data_ptr()

        r1 = lambda op: P(F1.rho[op][k1])
        r2 = lambda op: P(F2.rho[op][k2])
        j = PairJob()
        j.nI1, j.nJ1 = F1.n_states[k1[0]], F1.n_states[k1[1]]
        j.nI2, j.nJ2 = F2.n_states[k2[0]], F2.n_states[k2[1]]
        j.nuc = 0.0
        if d in (-2, 2):
            j.kind = PAIR_2MIN2PLS
            if d == -2:
                j.transposed, j.n1, j.n2, j.scale = 0, F1.dim, F2.dim, half
                j.R1[0], j.R2[0], j.V[0] = r1("cc"), r2("aa"), P(I.V(m1, m1, m2, m2))
            else:
                j.transposed, j.n1, j.n2, j.scale = 1, F2.dim, F1.dim, half
                j.R1[0], j.R2[0], j.V[0] = r2("cc"), r1("aa"), P(I.V(m2, m2, m1, m1))
        elif d in (-1, 1):
            j.kind = PAIR_1MIN1PLS
            if d == -1:
                j.transposed, j.n1, j.n2, j.scale = 0, F1.dim, F2.dim, half * sign2
                j.R1[0], j.R1[1] = r1("c"), r1("cca")
                j.R2[0], j.R2[1] = r2("a"), r2("caa")
                j.V[0], j.V[1] = P(I.V(m1, m1, m1, m2)), P(I.V(m1, m2, m2, m2))
                j.h[0] = P(I.h_sum((m1, m2, "1min"), [("T", (m1, m2)), ("U", (m1, m1, m2)), ("U", (m2, m1, m2))]))
            else:
                j.transposed, j.n1, j.n2, j.scale = 1, F2.dim, F1.dim, -half * sign2
                j.R1[0], j.R1[1] = r2("c"), r2("cca")
                j.R2[0], j.R2[1] = r1("a"), r1("caa")
                j.V[0], j.V[1] = P(I.V(m2, m2, m2, m1)), P(I.V(m2, m1, m1, m1))
                j.h[0] = P(I.h_sum((m2, m1, "1pls"), [("T", (m2, m1)), ("U", (m1, m2, m1)), ("U", (m2, m2, m1))]))
        else:
            j.kind, j.transposed, j.n1, j.n2, j.scale = PAIR_EXEX, 0, F1.dim, F2.dim, half
            j.R1[0], j.R2[0], j.V[0] = r1("ca"), r2("ca"), P(I.V(m1, m2, m1, m2))
            j.h[0], j.h[1] = P(I.U(m1, m2, m2)), P(I.U(m2, m1, m1))
            j.nuc = float(self.nuc[m1, m2])
        j.out, j.ldo, j.mul = P(out), int(ldo), int(mul)
        j.rows1, j.rows2 = P(rows[0]), P(rows[1])
        j.cols1, j.cols2 = P(cols[0]), P(cols[1])
        return j

    def pair_jobs(self, pairs, place):
        """(jobs, keep) for the 'self' (half weight) and 'source' charge
        combinations of the pairs (the outputs are completed as A + A^T);
        place(m1, m2, k1, k2) -> (out, ldo, rows, cols, mul)."""
        self._need()
        jobs, keep = [], []
        for (m1, m2) in pairs:
            for k1, k2 in self.pair_charge_combos(m1, m2):
                role = self.combo_role(k1, k2)
                if role == "derived":
                    continue
                out, ldo, rows, cols, mul = place(m1, m2, k1, k2)
                jobs.append(self._pair_job(m1, m2, k1, k2, 0.5 if role == "self" else 1.0,
                                           out, ldo, rows, cols, mul, keep))
        return jobs, keep

    def _dense_place(self, H):
        def place(m1, m2, k1, k2):
            F1, F2 = self.frags[m1], self.frags[m2]
            return (H[(m1, m2)], F1.d * F2.d, [F1.pos[k1[0]], F2.pos[k2[0]]],
                    [F1.pos[k1[1]], F2.pos[k2[1]]], F2.d)
        return place

    def pair_matrices(self, pairs):
        """{(m1, m2): H2[m1][m2]} for the pairs, all in one batched program
        (xr_pair_blocks); plain systems build each combination on its own."""
        if self.plain:
            return {(m1, m2): self.pair_matrix(m1, m2) for (m1, m2) in pairs}
        from .pairs import pair_blocks
        H = {(m1, m2): torch.empty(self.frags[m1].d * self.frags[m2].d, self.frags[m1].d * self.frags[m2].d,
                                   dtype=torch.float64, device=self.device) for (m1, m2) in pairs}
        jobs, keep = self.pair_jobs(pairs, self._dense_place(H))
        pair_blocks(jobs, list(H.values()), PRECISIONS[self.precision], self.device.index)
        return H

    def pair_plan(self, pairs, H=None):
        """A PairPlan writing the dense pair matrices of the pairs into H
        (allocated when not given): replays read the densities at their
        current addresses (rebuild them with reuse=True)."""
        from .pairs import PairPlan
        if H is None:
            H = {(m1, m2): torch.empty(self.frags[m1].d * self.frags[m2].d, self.frags[m1].d * self.frags[m2].d,
                                       dtype=torch.float64, device=self.device) for (m1, m2) in pairs}
        jobs, keep = self.pair_jobs(pairs, self._dense_place(H))
        return PairPlan(jobs, [H[p] for p in pairs], PRECISIONS[self.precision], self.device.index, keep), H

    def pair_matrix(self, m1, m2):
        """H2[m1][m2] over the product basis [(a, b) for a in states(m1) for b
        in states(m2)] = build_matrix_elements.dimer (build_H.py:56-102), one
        charge combination at a time."""
        F1, F2 = self.frags[m1], self.frags[m2]
        D = F1.d * F2.d
        out = torch.zeros(D, D, dtype=torch.float64, device=self.device)
        dims = [F1.d, F2.d]
        with precision(self.precision, self.device.index):
            self._pair_terms(m1, m2, lambda k1, k2: (out, D, [F1.pos[k1[0]], F2.pos[k2[0]]],
                                                     [F1.pos[k1[1]], F2.pos[k2[1]]], dims))
        context(self.device.index).call("xr_symmetrize", ptr(out), D)
        return out

    def _sector_place(self, m1, m2, segs, segq, blocks):
        """Placement into the fixed-charge blocks of pair (m1, m2): pair (a, b)
        is row off_q[a] + r(b) of block q = charge(a) + charge(b), r(b) being
        b's rank among the states of its charge and off_q[a] the number of
        charge-q pairs whose first state comes before a (blocks ordered by
        increasing a * d2 + b; segs[i] = (start, count) in that ordering)."""
        F1, F2 = self.frags[m1], self.frags[m2]
        cache = self._sector_idx.setdefault((m1, m2), {})
        if "layout" not in cache:
            ch1 = np.array([c for c, _ in F1.state_indices])
            ch2 = np.array([c for c, _ in F2.state_indices])
            off, rank2 = sector_layout(ch1, ch2, segq)
            pos1 = {c: F1.pos[c].cpu().numpy() for c in F1.charges}
            pos2 = {c: F2.pos[c].cpu().numpy() for c in F2.charges}
            cache["layout"] = (off, rank2, pos1, pos2)
        off, rank2, pos1, pos2 = cache["layout"]
        seg_of = {q: i for i, q in enumerate(segq)}
        dev = self.device

        def place(k1, k2):
            q = k1[0] + k2[0]
            i = seg_of[q]
            key = (k1, k2)
            if key not in cache:
                t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device=dev)
                cache[key] = ([t(off[q][pos1[k1[0]]]), t(rank2[pos2[k2[0]]])],
                              [t(off[q][pos1[k1[1]]]), t(rank2[pos2[k2[1]]])])
            rows, cols = cache[key]
            return blocks[i], segs[i][1], rows, cols
        return place

    def pair_sector_blocks(self, m1, m2, segs, segq):
        """H2[m1][m2] as its square blocks of fixed pair charge, without the
        dense pair matrix (layout of _sector_place), one charge combination
        at a time."""
        blocks = [torch.zeros(cnt, cnt, dtype=torch.float64, device=self.device) for _, cnt in segs]
        pl = self._sector_place(m1, m2, segs, segq, blocks)

        def place(k1, k2):
            out, ldo, rows, cols = pl(k1, k2)
            return out, ldo, rows, cols, [1, 1]

        with precision(self.precision, self.device.index):
            self._pair_terms(m1, m2, place)
        ctx = context(self.device.index)
        for blk, (_, cnt) in zip(blocks, segs):
            ctx.call("xr_symmetrize", ptr(blk), cnt)
        return blocks

    def pair_sector_blocks_all(self, layouts):
        """{(m1, m2): fixed-charge blocks} for layouts {(m1, m2): (segs,
        segq)}, every pair in one batched program (xr_pair_blocks)."""
        if self.plain:
            return {p: self.pair_sector_blocks(p[0], p[1], *lay) for p, lay in layouts.items()}
        from .pairs import pair_blocks
        blocks = {p: [torch.empty(cnt, cnt, dtype=torch.float64, device=self.device) for _, cnt in lay[0]]
                  for p, lay in layouts.items()}
        places = {p: self._sector_place(p[0], p[1], lay[0], lay[1], blocks[p]) for p, lay in layouts.items()}

        def place(m1, m2, k1, k2):
            out, ldo, rows, cols = places[(m1, m2)](k1, k2)
            return out, ldo, rows, cols, 1

        jobs, keep = self.pair_jobs(list(layouts), place)
        pair_blocks(jobs, [b for p in layouts for b in blocks[p]], PRECISIONS[self.precision], self.device.index)
        return blocks

    def trimer_matrix(self, m1, m2, m3):
        """H3[m1][m2][m3] = build_matrix_elements.trimer (build_H.py:103-188)."""
        self._need()
        ms = (m1, m2, m3)
        Fs = [self.frags[m] for m in ms]
        dims = [F.d for F in Fs]
        D = dims[0] * dims[1] * dims[2]
        out = torch.zeros(D, D, dtype=torch.float64, device=self.device)
        charges = [F.charges for F in Fs]
        for cb in [(a, b, c) for a in charges[0] for b in charges[1] for c in charges[2]]:
            for ck in [(a, b, c) for a in charges[0] for b in charges[1] for c in charges[2]]:
                if sum(cb) != sum(ck):
                    continue
                keys = list(zip(cb, ck))
                res = B.trimer_block(ms, [F.n_elec_ref for F in Fs], [F.rho for F in Fs], keys, self.ints)
                if res is None:
                    continue
                Hb, order, scale, extra = res
                fr = [(Fs[k].n_states[cb[k]], Fs[k].n_states[ck[k]]) for k in range(3)]
                npair = [fr[k][0] * fr[k][1] for k in range(3)]
                rows = [Fs[k].pos[cb[k]] for k in range(3)]
                cols = [Fs[k].pos[ck[k]] for k in range(3)]
                strides = [0, 0, 0]
                strides[order[0]] = npair[order[1]] * npair[order[2]]
                strides[order[1]] = npair[order[2]]
                strides[order[2]] = 1
                B.scatter(out, D, fr, rows, cols, dims, Hb, strides, scale=scale)
                if extra is not None:
                    spect, u, (y, z), sc = extra
                    if fr[spect][0] == fr[spect][1] and cb[spect] == ck[spect]:
                        st = [0, 0, 0]
                        st[y] = npair[z]
                        st[z] = 1
                        dg = [0, 0, 0]
                        dg[spect] = 1
                        B.scatter(out, D, fr, rows, cols, dims, u, st, diag=dg, scale=sc)
        return out

    # -- XR matrix ----------------------------------------------------------
    def blocks(self, with_trimers=False):
        with precision(self.precision, self.device.index):
            return self._blocks(with_trimers)

    def _blocks(self, with_trimers, pair_plan=None):
        """(H1, H2, H3); pair_plan: a PairPlan (pair_plan()) whose run
        gives the pair matrices instead of a fresh batched build."""
        nf = len(self.frags)
        H1 = [self.monomer_matrix(m) for m in range(nf)]
        H2 = [[None] * nf for _ in range(nf)]
        pairs = [(m, n) for m in range(nf) for n in range(m + 1, nf)]
        if pair_plan is not None:
            pm = dict(zip(pairs, pair_plan.run()))
        else:
            pm = self.pair_matrices(pairs)
        for (m, n), h in pm.items():
            H2[m][n] = h
        H3 = None
        if with_trimers:
            H3 = [[[None] * nf for _ in range(nf)] for _ in range(nf)]
            for a in range(nf):
                for b in range(a + 1, nf):
                    for c in range(b + 1, nf):
                        H3[a][b][c] = self.trimer_matrix(a, b, c)
        return H1, H2, H3

    def state_charges(self):
        """Per fragment, the charge of each state in state order."""
        return [np.array([chg for chg, _ in F.state_indices], dtype=np.int64) for F in self.frags]

    def operator(self, H=None):
        """Matrix-free XR operator (monomer and pair terms) over the blocks H,
        or, without H, with the pair terms built straight into the operator's
        fixed-charge blocks (no dense pair matrices)."""
        from .xr_operator import XROperator
        if H is None and self.plain:
            op = self.operator(H=self.blocks())
            op.grouped = False
            return op
        if H is None:
            nf = len(self.frags)
            with precision(self.precision, self.device.index):
                H1 = [self.monomer_matrix(m) for m in range(nf)]
                return XROperator(H1, {(m, n): None for m in range(nf) for n in range(m + 1, nf)},
                                  self.state_charges(), pairs_builder=self.pair_sector_blocks_all)
        H1, H2, H3 = H
        if H3 is not None:
            raise ValueError("the matrix-free operator carries monomer and pair terms only")
        nf = len(self.frags)
        pairs = {(m, n): H2[m][n] for m in range(nf) for n in range(m + 1, nf)}
        # pair_matrix returns exactly symmetric matrices in every precision
        return XROperator(H1, pairs, self.state_charges())

    def xr_matrix(self, with_trimers=False, spectator_sign=False, H=None):
        """Tensor-product XR matrix (all fragments need the same number of
        states).  spectator_sign adds the sign the reference assembly omits
        for one-electron transfers across spectator fragments."""
        H1, H2, H3 = H if H is not None else self.blocks(with_trimers)
        M = assemble(H1, H2, H3, [F.parity() for F in self.frags] if spectator_sign else None)
        if self.precision != "f64":
            # bra/ket-mirrored blocks come from different split products, so
            # the matrix is symmetric only to the split error; the eigensolver
            # gets its symmetric part (same v^T M v for every v)
            M = torch.add(M, M.T).mul_(0.5)
        return M


def assemble(H1, H2, H3=None, parities=None):
    """braket_loops on the device (general-XRCC/hamiltonian.py:82-83)."""
    nf = len(H1)
    spf = H1[0].shape[0]
    for h in H1:
        if h.shape[0] != spf:
            raise AssertionError("all fragments need the same number of states")
    dev = H1[0].device
    n = spf ** nf
    out = torch.empty(n, n, dtype=torch.float64, device=dev)
    p1 = (ctypes.c_void_p * nf)(*[h.contiguous().data_ptr() for h in H1])
    p2 = (ctypes.c_void_p * (nf * nf))(*[
        (H2[m][k].data_ptr() if (m < k and H2[m][k] is not None) else None)
        for m in range(nf) for k in range(nf)])
    p3 = None
    if H3 is not None:
        p3 = (ctypes.c_void_p * (nf ** 3))(*[
            (H3[a][b][c].data_ptr() if (a < b < c and H3[a][b][c] is not None) else None)
            for a in range(nf) for b in range(nf) for c in range(nf)])
    par = None
    if parities is not None:
        par = torch.cat([p.reshape(-1) for p in parities]).contiguous()
    context(dev.index).call("xr_braket_assemble", nf, spf, p1, p2, p3, ptr(par), ptr(out))
    return out
