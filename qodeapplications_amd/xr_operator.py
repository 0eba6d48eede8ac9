"""Matrix-free XR Hamiltonian on the tensor-product basis.

H = sum_m H1_m (acting on fragment axis m) + sum_{m<n} H2_mn (acting on the
axes m, n), the sums general-XRCC/hamiltonian.py:21-63 places element by
element into a dense matrix of side prod_m d_m.  Here a block of vectors is
a row-major tensor X[s_0]...[s_{N-1}][j] and H is applied without forming it:

* H1_m: one batched FP64 GEMM along axis m;
* H2_mn: xr_axes_pack gathers the (m, n) axes to the rows with the pair index
  sorted by the pair's total charge.  H conserves charge, so H2_mn in that
  order is block diagonal (one block per charge sum) and the product is one
  GEMM per block; xr_axes_pack adds the result back.  The blocks hold
  sum_s n_s^2 of the (sum_s n_s)^2 entries (19 / 81 for three equal charge
  blocks).

The same packing gives the reduced matrices the gradient needs:
D_m = sum_rest v v^T on fragment m, and the pair blocks G_mn of
sum_rest v v^T on (m, n), block diagonal in the same order.

The spectator-fragment sign of one-electron transfers across fragments
between m and n is not applied, as in the reference loops.
"""
import contextlib

import numpy as np
import torch

from ._lib import context, ptr
from .eigen import davidson, precond_divide, upload
from .gemm import gemm_f64, mm, syrk


DENSE_SECTOR = 2048


def _block_size(k, size):
    """Davidson block: small sectors are bound by the per-iteration host work,
    so they take 4 new directions per root and iteration; large ones are bound
    by the products and take one."""
    return k * (4 if size <= (1 << 20) else 1)


def _axis_apply(A, src, dst, pre, n, post, transA=0, beta=0.0):
    """dst[p, :, q] = op(A) src[p, :, q] (+ beta dst) for row-major (pre, n,
    post) views: one GEMM src @ op(A)^T when post == 1, else one batched GEMM."""
    if post == 1:
        mm(src.view(pre, n), A, transB=0 if transA else 1, out=dst.view(pre, n), beta=beta)
    else:
        S, D = src.view(pre, n, post), dst.view(pre, n, post)
        for b0 in range(0, pre, 65535):        # the batch is a grid dimension of at most 65535
            b1 = min(pre, b0 + 65535)
            gemm_f64(A.unsqueeze(0), S[b0:b1], D[b0:b1], transA=transA, beta=beta)


def _arr(vals, dtype):
    return np.ascontiguousarray(np.asarray(vals, dtype=dtype))


def _ptr(a):
    """Address of a numpy argument array (kept alive by the caller)."""
    return a.ctypes.data


_I32_CACHE = {}


def _i32(a, device):
    """Device int32 copy of an index array.  The index arrays depend only on
    the fragments' charge structure, so the copies are kept (keyed by
    content) and reused by the operators of later iterations instead of
    paying one synchronous host-to-device copy each."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    key = (str(device), a.size, hash(a.tobytes()))
    hit = _I32_CACHE.get(key)
    if hit is not None and np.array_equal(hit[0], a):
        return hit[1]
    t = upload(a, device)
    if len(_I32_CACHE) > 4096:
        _I32_CACHE.clear()
    _I32_CACHE[key] = (a, t)
    return t


BOUND_MAX_BLOCK = 4096     # pair blocks with more states get no certified bound
BOUND_STEPS = 30           # Lanczos steps per pair block
_U64 = 2.0 ** -53          # unit roundoff of FP64


def charge_sector_bound(lines, Q):
    """max over mu of f(mu) = sum_p min over (lam, c) in lines[p] of
    (lam + mu c), minus mu Q.

    lines[p] lists, for term p of a decomposition of H, the lower bound lam
    of its lowest eigenvalue on each block the sector of total charge Q
    reaches and that block's coefficient c of the charge multiplier mu.
    f(mu) bounds the sector's lowest eigenvalue from below for every mu; it
    is concave and piecewise linear, so its maximum lies at mu = 0 or where
    two lines of one term cross.  A term with no finite line makes the bound
    -inf."""
    if not lines:
        return float("-inf")
    for L in lines:
        if not L or not any(np.isfinite(lam) for lam, _ in L):
            return float("-inf")

    def f(mu):
        return sum(min(lam + mu * c for lam, c in L) for L in lines) - mu * Q

    cands = {0.0}
    for L in lines:
        for i in range(len(L)):
            for j in range(i + 1, len(L)):
                (l1, c1), (l2, c2) = L[i], L[j]
                if c1 != c2 and np.isfinite(l1) and np.isfinite(l2):
                    cands.add((l1 - l2) / (c2 - c1))
    return max(f(mu) for mu in cands)


_SIDE = {}                 # device -> (stream, library context) of the sector-bound work


def _side(device):
    """A second stream on the device (default priority) with its own library
    context, for the sector bounds computed beside the Davidson."""
    hit = _SIDE.get(device.index)
    if hit is None:
        from ._lib import Context
        st = torch.cuda.Stream(device=device)
        with torch.cuda.stream(st):
            hit = _SIDE[device.index] = (st, Context(device.index))
    return hit


class _SectorBoundJob:
    """The staged work behind XROperator.sector_bounds: the dressed pair
    blocks and their Lanczos passes; once those reach the host, the shifted
    Cholesky tests; once their outcomes arrive (and any retry's), the block
    and sector bounds.  With a side stream every stage is queued there with
    its own library context, after the caller's earlier work on its stream,
    and step() only advances past stages that have finished, so the stages
    run beside the caller's own GPU work."""

    def __init__(self, op, steps, max_block, shares, side=False):
        self.op, self.steps, self.shares = op, steps, shares
        self.done, self.out, self.lam = False, {}, {}
        op.block_bounds = {}
        nf = op.nf
        if nf < 3 or len(op.pairs) != nf * (nf - 1) // 2:
            self.done = True
            return
        self.wgt = 1.0 / (nf - 1)
        dev = op.device
        self.pairs = list(op.pairs.items())
        # pair-state digits (a, b) of every pair, one cached device array
        digs, self.doff, o = [], {}, 0
        for (m, n), P in self.pairs:
            order = P["order_np"]
            digs.append(np.stack([order // op.dims[n], order % op.dims[n]]).reshape(-1))
            self.doff[(m, n)] = o
            o += 2 * order.size
        self.dig = _i32(np.concatenate(digs), dev)
        self.items = sorted((cnt, m, n, i) for (m, n), P in self.pairs for i, (_, cnt) in enumerate(P["segs"])
                            if cnt <= max_block)
        if not self.items:
            self._finish(None)
            return
        self.stream, self.ctx = (None, op.ctx)
        if side:
            self.stream, self.ctx = _side(dev)
            self.stream.wait_stream(torch.cuda.current_stream(dev))
        nb = len(self.items)
        sizes = np.array([cnt * cnt for cnt, _, _, _ in self.items], dtype=np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)])
        with self._on():
            self.buf = torch.empty(int(offs[-1]), dtype=torch.float64, device=dev)
            base = self.buf.data_ptr()
            self.ns = _arr([cnt for cnt, _, _, _ in self.items], np.int64)
            self.ptrs = _arr([base + 8 * int(offs[j]) for j in range(nb)], np.uint64)
            for j in range(nb):
                self._dress(j)
            stats = torch.empty(2 * nb * steps + nb, dtype=torch.float64, device=dev)
            sp_ = stats.data_ptr()
            for c0 in range(0, nb, 64):
                c1 = min(nb, c0 + 64)
                self.ctx.call("xr_lanczos", c1 - c0, _ptr(self.ns[c0:c1]), _ptr(self.ptrs[c0:c1]),
                              _ptr(self.ns[c0:c1]), steps, sp_ + 8 * c0 * steps, sp_ + 8 * (nb + c0) * steps,
                              sp_ + 8 * (2 * nb * steps + c0))
            self.host = self._to_host(stats)
        self.stage = 1

    def _on(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    def _to_host(self, t):
        """Queues the copy of t into pinned host memory and the event that marks it."""
        h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        self.ev = torch.cuda.Event()
        self.ev.record()
        self.keep = t
        return h

    def _dress(self, j):
        """Item j's pair block with its monomer shares into its slot of buf."""
        cnt, m, n, i = self.items[j]
        op = self.op
        P = op.pairs[(m, n)]
        start = P["segs"][i][0]
        blk = P["blocks"][i]
        if blk.stride(1) != 1:
            raise ValueError("pair block rows must be contiguous")
        pa = self.dig.data_ptr() + 4 * (self.doff[(m, n)] + start)
        pb = pa + 4 * P["order_np"].size
        self.ctx.call("xr_pair_dress", cnt, blk.data_ptr(), blk.stride(0), pa, pb, op.H1[m].data_ptr(),
                      op.dims[m], op.H1[n].data_ptr(), op.dims[n], self.shares[(m, (m, n))],
                      self.shares[(n, (m, n))], int(self.ptrs[j]), cnt)

    def _cholesky(self, sel, shift):
        """Queues: shift[q] added to the diagonal of block sel[q], the
        in-place Cholesky tests of those blocks, and the copy of their info
        codes to the host."""
        ns, ptrs = _arr(self.ns[sel], np.int64), _arr(self.ptrs[sel], np.uint64)
        shift = _arr(shift, np.float64)
        with self._on():
            info = torch.empty(len(sel), dtype=torch.int32, device=self.op.device)
            for c0 in range(0, len(sel), 64):
                c1 = min(len(sel), c0 + 64)
                self.ctx.call("xr_add_diagonal", c1 - c0, _ptr(ns[c0:c1]), _ptr(ptrs[c0:c1]), _ptr(ns[c0:c1]),
                              _ptr(shift[c0:c1]))
                self.ctx.call("xr_cholesky_test", c1 - c0, _ptr(ns[c0:c1]), _ptr(ptrs[c0:c1]), _ptr(ns[c0:c1]),
                              info.data_ptr() + 4 * c0)
            return self._to_host(info)

    def step(self, block=False):
        """Advances through every stage whose work has finished (waiting for
        it when block); True once the bounds are final."""
        while not self.done:
            if not block and not self.ev.query():
                return False
            self.ev.synchronize()
            if self.stage == 1:
                self._estimates()
            elif self.stage == 2:
                self._outcomes()
            else:
                self.info[self.retry] = self.host.numpy()
                self._finish(self.info)
        return True

    def _estimates(self):
        """Lanczos tridiagonals -> theta, residual, the first shift t; queues
        the Cholesky tests."""
        nb, steps = len(self.items), self.steps
        h = self.host.numpy()
        A_, B_ = h[:nb * steps].reshape(nb, steps), h[nb * steps:2 * nb * steps].reshape(nb, steps)
        self.anorm = h[2 * nb * steps:].copy()
        theta, resid = np.full(nb, np.nan), np.full(nb, np.nan)
        for j in range(nb):
            a_, b_ = A_[j], B_[j]
            tiny = 1e-14 * max(self.anorm[j], 1e-300)
            kk = steps
            for q in range(steps):
                if not b_[q] > tiny:            # breakdown (or NaN): T of the first q + 1 steps
                    kk = q + 1
                    break
            T = np.diag(a_[:kk]) + np.diag(b_[:kk - 1], 1) + np.diag(b_[:kk - 1], -1)
            if np.all(np.isfinite(T)):
                w, y = np.linalg.eigh(T)
                theta[j], resid[j] = w[0], abs(b_[kk - 1] * y[kk - 1, 0])
        self.theta, self.resid = theta, resid
        self.ok = np.isfinite(theta) & np.isfinite(resid) & np.isfinite(self.anorm)
        self.t = np.where(self.ok, theta - np.maximum(2 * resid, 1e-3 * np.maximum(1.0, np.abs(theta))), 0.0)
        self.sel = np.flatnonzero(self.ok)
        self.info = np.ones(nb, dtype=np.int64)         # not certified
        if self.sel.size == 0:
            self._finish(self.info)
            return
        self.host = self._cholesky(self.sel, -self.t[self.sel])
        self.stage = 2

    def _outcomes(self):
        """First outcomes; a failed block is formed again (the factorisation
        overwrote it) and tested once more at a lower shift."""
        self.info[self.sel] = self.host.numpy()
        self.retry = np.array([j for j in self.sel if self.info[j] != 0], dtype=np.int64)
        if self.retry.size == 0:
            self._finish(self.info)
            return
        for j in self.retry:
            self.t[j] = self.theta[j] - max(8 * self.resid[j], 0.05 * max(1.0, abs(self.theta[j])))
        with self._on():
            for j in self.retry:
                self._dress(int(j))
        self.host = self._cholesky(self.retry, -self.t[self.retry])
        self.stage = 3

    def _finish(self, info):
        op = self.op
        lam = {}
        if info is not None:
            for j, (cnt, m, n, i) in enumerate(self.items):
                if self.ok[j] and info[j] == 0:
                    slack = 4.0 * (cnt + 1) * cnt * _U64 * (self.anorm[j] + abs(self.t[j]))
                    lam[(m, n, op.pairs[(m, n)]["segq"][i])] = float(self.t[j] - slack)
        self.buf = self.keep = None
        op.block_bounds = lam
        cs = [sorted(set(int(x) for x in c)) for c in op.charges]
        rest = {}
        for (m, n), _ in self.pairs:
            tot = {0}
            for f in range(op.nf):
                if f not in (m, n):
                    tot = {x + y for x in tot for y in cs[f]}
            rest[(m, n)] = tot
        self.out = {}
        for Q in op.total_charges():
            lines = [[(lam.get((m, n, s), float("-inf")), s * self.wgt) for s in P["segq"] if Q - s in rest[(m, n)]]
                     for (m, n), P in self.pairs]
            self.out[Q] = charge_sector_bound(lines, Q)
        self.done = True


class XROperator:
    """H1: list of monomer matrices (d_m x d_m); H2: {(m, n): dense pair
    matrix (d_m d_n x d_m d_n)} for m < n; charges: per fragment, the charge
    of each state in state order."""

    def __init__(self, H1, H2, charges, symmetrize=False, pair_builder=None, pairs_builder=None):
        self.H1 = [h.contiguous() for h in H1]
        self.dims = [int(h.shape[0]) for h in H1]
        self.nf = len(self.dims)
        self.n = int(np.prod(self.dims))
        self.device = H1[0].device
        self.ctx = context(self.device.index)
        self.charges = [np.asarray(c, dtype=np.int64) for c in charges]
        self.pairs = {}
        self._sq, self._pair_diag, self._sqidx = {}, {}, {}
        self.found = None
        self.log = None          # set to a list to record the per-sector solves
        self.grouped = True      # pair products of a sector in one grouped GEMM (else one launch each)
        pending = {}
        for (m, n), H in H2.items():
            if H is not None:
                self.pairs[(m, n)] = self._pair_blocks(m, n, H, symmetrize)
            elif pairs_builder is not None:
                pending[(m, n)] = self._order(m, n)
            elif pair_builder is not None:
                # pair_builder(m, n, segs, segq) returns the fixed-charge
                # blocks in the layout _pair_blocks cuts out of a dense matrix
                order, inv, segs, segq = self._order(m, n)
                self.pairs[(m, n)] = self._pair_meta(order, inv, segs, segq, pair_builder(m, n, segs, segq))
        if pending:
            # pairs_builder({(m, n): (segs, segq)}) returns those blocks of
            # every pair at once
            built = pairs_builder({p: (o[2], o[3]) for p, o in pending.items()})
            for p, (order, inv, segs, segq) in pending.items():
                self.pairs[p] = self._pair_meta(order, inv, segs, segq, built[p])

    # -- setup ----------------------------------------------------------------
    def _order(self, m, n):
        q = (self.charges[m][:, None] + self.charges[n][None, :]).ravel()
        order = np.argsort(q, kind="stable")
        inv = np.empty_like(order)
        inv[order] = np.arange(order.size)
        segs, segq = [], []
        qs = q[order]
        start = 0
        for i in range(1, qs.size + 1):
            if i == qs.size or qs[i] != qs[start]:
                segs.append((start, i - start))
                segq.append(int(qs[start]))
                start = i
        return order, inv, segs, segq

    def _pair_blocks(self, m, n, H, symmetrize):
        D = self.dims[m] * self.dims[n]
        if tuple(H.shape) != (D, D):
            raise ValueError(f"pair matrix ({m}, {n}) must be {D} x {D}")
        order, inv, segs, segq = self._order(m, n)
        H = H.contiguous()
        blocks = []
        for start, cnt in segs:
            rows = _i32(order[start:start + cnt], self.device)
            # tmp[r, x] = H[x, rows[r]]; blk[r, x] = tmp[x, rows[r]] = H[rows[r], rows[x]]
            tmp = torch.empty(cnt, D, dtype=torch.float64, device=self.device)
            self._pack([D, D], 1, 1, -1, rows, cnt, H, tmp, 0)
            blk = torch.empty(cnt, cnt, dtype=torch.float64, device=self.device)
            self._pack([cnt, D], 1, 1, -1, rows, cnt, tmp, blk, 0)
            del tmp
            if symmetrize:
                blk = torch.add(blk, blk.T).mul_(0.5)
            blocks.append(blk)
        return self._pair_meta(order, inv, segs, segq, blocks)

    def _pair_meta(self, order, inv, segs, segq, blocks):
        return {"order": _i32(order, self.device), "inv": _i32(inv, self.device), "segs": segs,
                "segq": segq, "blocks": blocks, "inv_np": inv, "order_np": order}

    def extract(self, m, n, rows, cols):
        """H2_mn[rows][:, cols] for pair indices a * d_n + b of one charge sum."""
        P = self.pairs[(m, n)]
        r, c = P["inv_np"][rows], P["inv_np"][cols]
        for i, (start, cnt) in enumerate(P["segs"]):
            if start <= r.min() and r.max() < start + cnt:
                break
        else:
            raise ValueError("rows span several charge sums")
        if c.min() < start or c.max() >= start + cnt:
            raise ValueError("rows and columns have different charge sums")
        blk = P["blocks"][i]
        lc = _i32(c - start, self.device)
        lr = _i32(r - start, self.device)
        tmp = torch.empty(c.size, cnt, dtype=torch.float64, device=self.device)
        self._pack([cnt, cnt], 1, 1, -1, lc, c.size, blk, tmp, 0)       # tmp[x, y] = blk[y, lc[x]]
        out = torch.empty(r.size, c.size, dtype=torch.float64, device=self.device)
        self._pack([c.size, cnt], 1, 1, -1, lr, r.size, tmp, out, 0)    # out[x, y] = blk[lr[x], lc[y]]
        return out

    def _monomer_eig(self, h1):
        """Eigenpairs of every (fragment, charge) monomer block, computed once
        per operator on the host (blocks of at most a few hundred states; a
        device solver would launch thousands of tiny kernels)."""
        if getattr(self, "_eig", None) is None:
            # all blocks in one transfer
            keys = [(m, c) for m, blocks in enumerate(h1) for c in blocks]
            flat = torch.cat([h1[m][c].reshape(-1) for m, c in keys]).cpu().numpy()
            self._eig = [{} for _ in h1]
            self._eig_low = {}         # (m, c) -> lowest eigenvalue, on the host
            off = 0
            for m, c in keys:
                n = h1[m][c].shape[0]
                hh = flat[off:off + n * n].reshape(n, n)
                off += n * n
                w, U = np.linalg.eigh(0.5 * (hh + hh.T))
                self._eig[m][c] = (upload(w, self.device), upload(U, self.device))
                self._eig_low[(m, c)] = float(w[0])
        return self._eig

    def _dressed_eig(self, sp, eig0):
        """Per fragment m and charge c: (diag(U^T h_m U), U), U the
        eigenvectors of the monomer block dressed with the mean field of
        every partner n in its lowest monomer state |0_n>,
        h~_m = h_m + sum_n <0_n| H2_mn |0_n>.  The preconditioner's product
        basis built from these captures the average pair coupling: on the
        4-fragment chain it converges in 17 Davidson iterations instead of
        63 with the bare monomer eigenbasis.  Computed once per operator
        (sp: any sector space, for the square pair blocks)."""
        if getattr(self, "_deig", None) is not None:
            return self._deig
        nf = self.nf
        ground = []                   # (charge, |0_n><0_n|) of each fragment's lowest monomer state
        for n in range(nf):
            c0 = min(eig0[n], key=lambda c: self._eig_low[(n, c)])
            ground.append((c0, eig0[n][c0][1][:, 0].contiguous()))
        # the dressed blocks of every (m, c) are formed on the device and
        # reach the host in one transfer
        dressed = []

        def gemm(M, N, K, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch):
            self.ctx.call("xr_gemm_f64", 0, 0, M, N, K, 1.0, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch)

        for m in range(nf):
            for c, h in sp.h1[m].items():
                ht = h.clone()
                for (a, b) in self.pairs:
                    if m not in (a, b):
                        continue
                    n = b if a == m else a
                    cn, u = ground[n]
                    ca, cb = (c, cn) if a == m else (cn, c)
                    Hss, parts = sp.sq[(a, b, ca + cb)]
                    o, D = [(o, D) for c1, c2, o, D in parts if (c1, c2) == (ca, cb)][0]
                    da, db = sp.blocks_of[a][ca][1], sp.blocks_of[b][cb][1]
                    # ht += (<u| on the partner) M (|u> on the partner), M the
                    # pair block over (a, b) states: two batched products on
                    # the strided block, the rows of M the batch
                    blk, ld = Hss.data_ptr() + 8 * (o * Hss.shape[1] + o), Hss.shape[1]
                    if a == m:         # rows (i, x): W[(i,x), j] = sum_y M[(i,x), (j,y)] u_y
                        W = torch.empty(D, da, dtype=torch.float64, device=self.device)
                        gemm(da, 1, db, blk, db, ld, u.data_ptr(), 1, 0, 0.0, W.data_ptr(), 1, da, D)
                        # ht[i, :] += sum_x u_x W[(i,x), :]
                        gemm(1, da, db, u.data_ptr(), db, 0, W.data_ptr(), da, db * da, 1.0, ht.data_ptr(), da,
                             da, da)
                    else:              # rows (x, i): W[(x,i), j] = sum_y u_y M[(x,i), (y,j)]
                        W = torch.empty(D, db, dtype=torch.float64, device=self.device)
                        gemm(1, db, da, u.data_ptr(), da, 0, blk, db, ld, 0.0, W.data_ptr(), db, db, D)
                        # ht += sum_x u_x W[(x, :), :]
                        gemm(1, db * db, da, u.data_ptr(), da, 0, W.data_ptr(), db * db, 0, 1.0, ht.data_ptr(),
                             db * db, 0, 1)
                dressed.append((m, c, h, ht))
        flat = torch.cat([ht.reshape(-1) for _, _, _, ht in dressed]).cpu().numpy()
        self._deig = [{} for _ in range(nf)]
        off = 0
        for m, c, h, ht in dressed:
            n = ht.shape[0]
            hh = flat[off:off + n * n].reshape(n, n)
            off += n * n
            _, U = np.linalg.eigh(0.5 * (hh + hh.T))
            U = upload(U, self.device)
            w = torch.empty(n, dtype=torch.float64, device=self.device)      # diag(U^T h U)
            self.ctx.call("xr_column_dots", n, n, U.data_ptr(), n, mm(h, U).data_ptr(), n, w.data_ptr())
            self._deig[m][c] = (w, U)
        return self._deig

    def _pack(self, dims, ncol, f0, f1, rowmap, nrows, full, packed, mode, alpha=1.0, ld=0):
        import ctypes
        d = (ctypes.c_int64 * len(dims))(*[int(x) for x in dims])
        self.ctx.call("xr_axes_pack", len(dims), d, int(ncol), int(f0), int(f1),
                      None if rowmap is None else ptr(rowmap), int(nrows), ptr(full), ptr(packed),
                      int(ld), int(mode), float(alpha))

    # -- products -------------------------------------------------------------
    def matvec(self, X, Y):
        """Y = H X for X, Y of shape (n, b), contiguous."""
        b = X.shape[1]
        dims = self.dims
        for m, h in enumerate(self.H1):
            pre = int(np.prod(dims[:m]))
            post = int(np.prod(dims[m + 1:])) * b
            X3 = X.view(pre, dims[m], post)
            Y3 = Y.view(pre, dims[m], post)
            gemm_f64(h.unsqueeze(0), X3, Y3, beta=0.0 if m == 0 else 1.0)
        for (m, n), P in self.pairs.items():
            D = dims[m] * dims[n]
            R = self.n // D * b
            Xp = torch.empty(D, R, dtype=torch.float64, device=self.device)
            self._pack(dims, b, m, n, P["order"], D, X, Xp, 0)
            Yp = torch.empty_like(Xp)
            for (start, cnt), blk in zip(P["segs"], P["blocks"]):
                mm(blk, Xp[start:start + cnt], out=Yp[start:start + cnt])
            self._pack(dims, b, m, n, P["order"], D, Y, Yp, 1, 1.0)
        return Y

    def dense(self):
        """The full matrix (small systems only): H applied to the identity."""
        E = torch.eye(self.n, dtype=torch.float64, device=self.device)
        out = torch.empty_like(E)
        self.matvec(E, out)
        return out

    # -- eigenpairs -----------------------------------------------------------
    def _sector_sizes(self):
        """{Q: number of product states of total charge Q}."""
        sizes = {0: 1}
        for c in self.charges:
            vals, cnts = np.unique(c, return_counts=True)
            nxt = {}
            for t, n in sizes.items():
                for v, cnt in zip(vals, cnts):
                    nxt[t + int(v)] = nxt.get(t + int(v), 0) + n * int(cnt)
            sizes = nxt
        return sizes

    def sector_space(self, Q):
        """The states of total charge Q in the compact sector layout."""
        return SectorSpace(self, Q)

    def total_charges(self):
        qs = [sorted(set(int(x) for x in c)) for c in self.charges]
        tot = {0}
        for q in qs:
            tot = {t + x for t in tot for x in q}
        return sorted(tot)

    def sector_bounds(self, steps=BOUND_STEPS, max_block=BOUND_MAX_BLOCK):
        """{Q: a certified lower bound of the lowest eigenvalue in the sector
        of total charge Q}, from the pair decomposition of H ({} unless
        there are at least 3 fragments and every pair has its term).

        Every fragment lies in the nf - 1 pairs (m, n), so
        H = sum_(m<n) h_mn with h_mn = H2_mn + w_m,mn H1_m + w_n,mn H1_n for
        any shares w >= 0 with sum_n w_m,mn = 1 (monomer_shares), and the
        total charge N = sum_m n_m is Q on the sector.  For a unit vector v
        of the sector and any mu,
        <v|H|v> = -mu Q + sum_(m<n) <v| h_mn + mu (n_m + n_n) / (nf - 1) |v>.
        Term (m, n) conserves the pair charge s = n_m + n_n, and on the
        sector s takes only the values the other fragments' charges complete
        to Q, so the term is at least the lowest of
        lambda_mn(s) + mu s / (nf - 1) over those s, lambda_mn(s) the lowest
        eigenvalue of h_mn on its block of charge sum s.  The bound is the
        maximum over mu (charge_sector_bound).

        lambda_mn(s) is certified from below: xr_pair_dress forms the block
        (symmetric parts), xr_lanczos gives an estimate theta with residual
        r, and t = theta - max(2 r, 1e-3 max(1, |theta|)) stands when the
        Cholesky factorisation of the block minus t I succeeds
        (xr_cholesky_test); lambda >= t minus the factorisation's
        backward-error allowance 4 (n + 1) n u (|A|_inf + |t|).  A failed
        factorisation is retried once at theta - max(8 r, 0.05 max(1,
        |theta|)); then, and for blocks of more than max_block states, the
        block's bound is -inf.  The certified block bounds are kept in
        self.block_bounds ((m, n, s) -> bound)."""
        sq = torch.empty(max(1, self._n_blocks()), dtype=torch.float64, device=self.device)
        self._block_sqnorms(sq)
        job = _SectorBoundJob(self, steps, max_block, self.monomer_shares_from(sq.cpu().numpy()))
        job.step(block=True)
        return job.out

    def _n_blocks(self):
        return sum(len(P["blocks"]) for P in self.pairs.values())

    def _block_sqnorms(self, out):
        """Queues out[j] = the squared Frobenius norm of pair block j (pairs
        in order, blocks in charge order; xr_column_norms on the block as one
        contiguous column)."""
        j = 0
        for P in self.pairs.values():
            for blk in P["blocks"]:
                n2 = blk.numel()
                if not blk.is_contiguous():
                    raise ValueError("pair blocks must be contiguous")
                self.ctx.call("xr_column_norms", n2, 1, blk.data_ptr(), 1, out.data_ptr() + 8 * j, 1)
                j += 1

    def monomer_shares_from(self, sq):
        """{(m, (m, n)): w}: fragment m's monomer term is shared over its
        pairs in proportion to the pairs' Frobenius norms (sq: the squared
        norms of the blocks, as _block_sqnorms orders them), uniform where
        every pair of m is zero.  Any shares give a valid bound; weighting by
        the coupling strength makes the strongly coupled pairs pay more of
        the monomer excitation energy their coupling exploits (on the
        4 x 13 chain it certifies one more sector than uniform shares)."""
        norm, j = {}, 0
        for p, P in self.pairs.items():
            k = len(P["blocks"])
            norm[p] = float(np.sqrt(max(0.0, float(np.sum(sq[j:j + k])))))
            j += k
        shares = {}
        for m in range(self.nf):
            ps = [p for p in self.pairs if m in p]
            tot = sum(norm[p] for p in ps)
            for p in ps:
                shares[(m, p)] = norm[p] / tot if tot > 0 and np.isfinite(tot) else 1.0 / len(ps)
        self.monomer_shares = shares
        return shares

    def lowest(self, k=1, tol=1e-9, max_iter=400, block=None, charges=None, chunk=4, margin=4.0,
               screen=True, verify=True, verify_tol=1e-4, bound=True):
        """(values[k], vectors[n, k]) of the lowest roots over all total
        charges (or the given ones), in the product layout.

        H conserves the total charge, so each charge sector is solved on its
        own compact layout (block size k unless given).  Small sectors are
        diagonalised densely.  The others are iterated in turns of `chunk`
        Davidson iterations, in the order of their lowest diagonal energy in
        the monomer eigenbasis, each resuming its own subspace.

        screen: every Ritz value is an upper bound to an eigenvalue of its
        sector, so the k-th lowest of all current Ritz values bounds the k-th
        lowest eigenvalue from above; a sector whose Ritz values minus
        `margin` times their residual norms all lie above that bound is set
        aside early (early Ritz values of a strongly correlated sector can
        lie well above its lowest eigenvalue: a margin of 2 set aside the
        true ground-state sector of a 3-fragment test chain after 3
        iterations at residual 0.55 and 2.1 Eh above the root).  Such a test
        bounds *some* eigenvalue of the sector, not its lowest one, so it
        only orders the work:

        verify (default): once the other sectors are converged, the lowest
        root of every set-aside sector is converged (resuming its subspace)
        to a residual of verify_tol * max(1, |value|); a sector whose
        converged lowest Ritz value minus its residual norm is not above the
        k-th root found so far is then converged to `tol` and competes like
        the others.  Every sector's lowest root is thus a converged Davidson
        root, the same standard as the root returned.  verify=False keeps
        the early drop final (faster, heuristic).  screen=False converges
        every sector to `tol` without setting any aside.

        bound (with screen, default): every sector first gets a certified
        lower bound of its lowest eigenvalue (sector_bounds); a sector whose
        bound lies above the k-th lowest Ritz value at hand (an upper bound
        of the k-th lowest eigenvalue) is set aside for good, with no
        further iteration.  The bounds are kept in self.bounds.

        self.log receives (Q, size, Ritz values, residual, iterations, how)
        per sector when it is a list; how is 'converged', 'dense',
        'verified-above', 'bounded-above' (the values are then [the bound])
        or 'dropped'."""
        found = []                    # converged: (value, sector space, vector)
        active = []
        iterative = []
        set_aside = []
        qs = charges if charges is not None else self.total_charges()
        # the sector bounds are computed on a side stream while the sector of
        # lowest diagonal energy (the likeliest home of the root) iterates
        self.bounds = {}
        job = None
        sizes = self._sector_sizes()
        bounding = screen and bound and sum(1 for Q in qs if sizes.get(Q, 0) > DENSE_SECTOR) > 1
        for Q in qs:
            sp = self.sector_space(Q)
            kk = min(k, sp.size)
            if sp.size <= DENSE_SECTOR:
                M = torch.empty(sp.size, sp.size, dtype=torch.float64, device=self.device)
                sp.matvec(torch.eye(sp.size, dtype=torch.float64, device=self.device), M)
                w, v = torch.linalg.eigh(0.5 * (M + M.T))
                found += [(float(w[i]), sp, v[:, i]) for i in range(kk)]
                if self.log is not None:
                    self.log.append((sp.Q, sp.size, [float(w[i]) for i in range(kk)], 0.0, 0, "dense"))
            else:
                iterative.append((sp, kk))
        if iterative:                 # the sectors' lowest diagonal energies (and the pair blocks' norms) in one transfer
            parts = [torch.stack([sp.lam.min() for sp, _ in iterative])]
            if bounding:
                sq = torch.empty(max(1, self._n_blocks()), dtype=torch.float64, device=self.device)
                self._block_sqnorms(sq)
                parts.append(sq)
            h = torch.cat(parts).cpu().numpy()
            lams = h[:len(iterative)]
            if bounding:
                job = _SectorBoundJob(self, BOUND_STEPS, BOUND_MAX_BLOCK,
                                      self.monomer_shares_from(h[len(iterative):]), side=True)
            for (sp, kk), lam in zip(iterative, lams):
                active.append({"sp": sp, "kk": kk, "b": block or _block_size(kk, sp.size), "st": {},
                               "w": None, "iters": 0, "lam": float(lam)})
        active.sort(key=lambda a: a["lam"])

        def upper(pending=True):
            vals = [f[0] for f in found]
            if pending:
                vals += [float(x) for a in active if a["w"] is not None for x in a["w"]]
            vals.sort()
            return vals[k - 1] if len(vals) >= k else None

        def record(a, inf, how):
            if self.log is not None:
                self.log.append((a["sp"].Q, a["sp"].size, a["w"], float(np.max(inf.get("rn", [0.0]))),
                                 a["iters"], how))

        def settle(block):
            """Takes the bounds once the job is done (waiting for it when block)."""
            nonlocal job
            if job is not None and job.step(block):
                self.bounds = job.out
                job = None
            return job is None

        def bounded(a, ub):
            """The sector's certified bound lies above ub: it is set aside."""
            lb = self.bounds.get(a["sp"].Q)
            if lb is None or ub is None or not lb > ub + 1e-9 * max(1.0, abs(ub)):
                return False
            a["st"].clear()
            a["w"] = [lb]
            record(a, {}, "bounded-above")
            return True

        def run(a):
            """One turn of `chunk` Davidson iterations on sector a."""
            sp, inf = a["sp"], {}
            w, v = davidson(sp.matvec, None, sp.size, a["kk"], tol, min(chunk, max_iter - a["iters"]), a["b"],
                            self.device, precond=sp.precondition,
                            start=sp.guess(a["b"]) if a["w"] is None else None, strict=False,
                            state=a["st"], stop_above=upper() if screen else None, stop_margin=margin,
                            info=inf)
            a["w"] = [float(x) for x in w]
            a["iters"] += inf.get("iters", 0)
            done = inf.get("abandoned", False) or inf.get("converged", False)
            if not done and a["iters"] >= max_iter:
                raise RuntimeError(f"Davidson did not converge in {max_iter} iterations in the sector of "
                                   f"total charge {sp.Q} (residual {float(np.max(inf['rn'])):.2e})")
            if done:
                active.remove(a)
                if inf.get("converged", False):
                    a["st"].clear()             # free the subspace
                    found.extend((float(w[i]), sp, v[:, i]) for i in range(w.shape[0]))
                    record(a, inf, "converged")
                elif verify:
                    set_aside.append(a)         # keeps its subspace for the verification
                else:
                    a["st"].clear()
                    record(a, inf, "dropped")

        first = active[0] if active else None
        while active:
            if not settle(False) and first in active:
                run(first)
                continue
            settle(True)
            for a in list(active):
                if bounded(a, upper()):
                    active.remove(a)
                    continue
                run(a)
        settle(True)
        for a in set_aside:
            if bounded(a, upper(pending=False)):
                continue
            sp, inf = a["sp"], {}
            w, v = davidson(sp.matvec, None, sp.size, 1, verify_tol, max(max_iter - a["iters"], 1), a["b"],
                            self.device, precond=sp.precondition, state=a["st"], info=inf)
            a["iters"] += inf.get("iters", 0)
            a["w"] = [float(x) for x in w]
            bound = upper(pending=False)
            if bound is None or float(w[0]) - float(np.max(inf.get("rn", [0.0]))) <= bound:
                inf = {}
                w, v = davidson(sp.matvec, None, sp.size, a["kk"], tol, max(max_iter - a["iters"], 1), a["b"],
                                self.device, precond=sp.precondition, state=a["st"], info=inf)
                a["iters"] += inf.get("iters", 0)
                a["w"] = [float(x) for x in w]
                found.extend((float(w[i]), sp, v[:, i]) for i in range(w.shape[0]))
                record(a, inf, "converged")
            else:
                record(a, inf, "verified-above")
            a["st"].clear()
        found.sort(key=lambda t: t[0])
        found = found[:k]
        self.found = [(sp, x) for _, sp, x in found]     # the roots in their sector layouts
        vals = torch.tensor([f[0] for f in found], dtype=torch.float64, device=self.device)
        vecs = torch.zeros(self.n, len(found), dtype=torch.float64, device=self.device)
        for i, (_, sp, x) in enumerate(found):
            sp.to_product(x.contiguous(), vecs[:, i])
        return vals, vecs

    # -- reduced matrices for the gradient ------------------------------------
    def monomer_density(self, v, m):
        """D_m[a, a'] = sum over the other fragments of v[.., a, ..] v[.., a', ..]."""
        d = self.dims[m]
        P = torch.empty(d, self.n // d, dtype=torch.float64, device=self.device)
        self._pack(self.dims, 1, m, -1, None, d, v, P, 0)
        return mm(P, P, transB=1)

    def pair_density(self, v, m, n, sector=None):
        """Diagonal blocks of G_mn = sum over the other fragments of v v^T on
        the (m, n) pair index, in the charge-sorted order of the pair blocks.
        sector = (sector space, v in its layout): the products then run over
        the sector's tuples only (v vanishes outside its charge sector)."""
        if sector is not None:
            return sector[0].pair_density(sector[1], m, n)
        P = self.pairs[(m, n)]
        D = self.dims[m] * self.dims[n]
        Vp = torch.empty(D, self.n // D, dtype=torch.float64, device=self.device)
        self._pack(self.dims, 1, m, n, P["order"], D, v, Vp, 0)
        blocks = []
        for start, cnt in P["segs"]:
            S = Vp[start:start + cnt]
            blocks.append(syrk(S))
        return blocks

    def pair_gather(self, G_blocks, m, n, pos, k1, k2, swapped):
        """dE/dH of the pair block with charges k1 = (bra, ket) on m, k2 on n,
        in the layout of the block's bilinear terms ((I1 J1) x (I2 J2), or
        transposed when swapped)."""
        P = self.pairs[(m, n)]
        i = P["segq"].index(k1[0] + k2[0])
        start, cnt = P["segs"][i]
        (r1, c1), (r2, c2) = (pos[0][k1[0]], pos[0][k1[1]]), (pos[1][k2[0]], pos[1][k2[1]])
        nI1, nJ1, nI2, nJ2 = r1.numel(), c1.numel(), r2.numel(), c2.numel()
        out = torch.empty(nI1 * nJ1 * nI2 * nJ2, dtype=torch.float64, device=self.device)
        self.ctx.call("xr_pair_gather", ptr(G_blocks[i]), cnt, self.dims[n], start, ptr(P["inv"]),
                      ptr(r1), ptr(c1), ptr(r2), ptr(c2), nI1, nJ1, nI2, nJ2, int(swapped), ptr(out))
        return out.view(nI2 * nJ2, nI1 * nJ1) if swapped else out.view(nI1 * nJ1, nI2 * nJ2)


class SectorSpace:
    """States of one total charge Q in the compact sector layout: the
    fragment-charge tuples c = (c_0, ..., c_{N-1}) with sum Q, each a dense
    row-major tensor of shape (n_{0,c_0}, ..., n_{N-1,c_{N-1}}), concatenated.
    H conserves Q, so the Davidson iteration of one sector never leaves it;
    monomer terms act inside a tuple, pair terms move the (m, n) charges of a
    tuple with their sum fixed."""

    def __init__(self, op, Q):
        self.op = op
        self.Q = Q
        dev = op.device
        nf = op.nf
        self.blocks_of = []          # per fragment: {charge: (start, count)} in state order
        for m in range(nf):
            q = op.charges[m]
            d = {}
            for c in sorted(set(int(x) for x in q)):
                idx = np.nonzero(q == c)[0]
                if idx.size and not np.all(np.diff(idx) == 1):
                    raise ValueError("the states of one charge must be contiguous")
                d[c] = (int(idx[0]), int(idx.size))
            self.blocks_of.append(d)
        charges = [sorted(d) for d in self.blocks_of]
        tuples = [()]
        for m in range(nf):
            tuples = [t + (c,) for t in tuples for c in charges[m]]
        self.sectors = [t for t in tuples if sum(t) == Q]
        self.index = {t: i for i, t in enumerate(self.sectors)}
        self.shapes = [[self.blocks_of[m][t[m]][1] for m in range(nf)] for t in self.sectors]
        self.sizes = [int(np.prod(sh)) for sh in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.size = int(self.offsets[-1])
        # monomer blocks per (fragment, charge) and their eigenbases
        self.h1 = [{c: op.H1[m][st:st + n, st:st + n].contiguous() for c, (st, n) in self.blocks_of[m].items()}
                   for m in range(nf)]
        self.eig = op._monomer_eig(self.h1)
        # pair terms: H2_mn conserves c_m + c_n, so for each pair and charge
        # sum s the block between all (c_m, c_n) with that sum is one square
        # matrix (rows and columns ordered by c_m, then by pair index); it is
        # shared by the sectors of every total charge
        self._plans = {}
        self.sq = op._sq
        for (m, n) in op.pairs:
            for s in sorted({cm + cn for cm in charges[m] for cn in charges[n]}):
                if (m, n, s) in self.sq:
                    continue
                parts, idx, off = [], [], 0
                for cm in charges[m]:
                    cn = s - cm
                    if cn in self.blocks_of[n]:
                        D = self.blocks_of[m][cm][1] * self.blocks_of[n][cn][1]
                        parts.append((cm, cn, off, D))
                        idx.append(self._pair_index(m, n, cm, cn))
                        off += D
                R = np.concatenate(idx)
                # states grouped by ascending charge: the fixed-charge pair
                # block is already in this order and is used as it is
                Pp = op.pairs[(m, n)]
                blk = None
                if s in Pp["segq"]:
                    i = Pp["segq"].index(s)
                    start, cnt = Pp["segs"][i]
                    if cnt == R.size and np.array_equal(R, Pp["order_np"][start:start + cnt]):
                        blk = Pp["blocks"][i]
                self.sq[(m, n, s)] = (op.extract(m, n, R, R) if blk is None else blk, parts)
                op._sqidx[(m, n, s)] = R
        # preconditioner basis: monomer blocks dressed with the partners' mean
        # field; then the diagonal of each charge-conserving pair block in it,
        # diag((Um x Un)^T H2[(cm,cn),(cm,cn)] (Um x Un))
        self.eig = op._dressed_eig(self, self.eig)
        self.pair_diag = op._pair_diag
        for (m, n, s), (Hss, parts) in self.sq.items():
            for cm, cn, off, D in parts:
                if (m, n, cm, cn) in self.pair_diag:
                    continue
                Dm, Dn = self.blocks_of[m][cm][1], self.blocks_of[n][cn][1]
                Hself = Hss[off:off + D, off:off + D]
                K = torch.kron(self.eig[m][cm][1], self.eig[n][cn][1]).contiguous()
                HK = mm(Hself, K)
                pdv = torch.empty(D, dtype=torch.float64, device=dev)
                op.ctx.call("xr_column_dots", D, D, K.data_ptr(), D, HK.data_ptr(), D, pdv.data_ptr())
                self.pair_diag[(m, n, cm, cn)] = pdv.view(Dm, Dn)
        # lam, tuple by tuple in the library: monomer then pair terms (all
        # pairs m < n in order; a pair without a term adds zeros)
        zero = {}
        pd = []
        for t, sh in zip(self.sectors, self.shapes):
            for m in range(nf):
                for n in range(m + 1, nf):
                    if (m, n) in op.pairs:
                        pd.append(self.pair_diag[(m, n, t[m], t[n])].data_ptr())
                    else:
                        z = zero.setdefault((sh[m], sh[n]), torch.zeros(sh[m], sh[n], dtype=torch.float64,
                                                                          device=dev))
                        pd.append(z.data_ptr())
        dims = _arr(self.shapes, np.int64).reshape(-1)
        offs = _arr(self.offsets[:-1], np.int64)
        e = _arr([self.eig[m][t[m]][0].data_ptr() for t in self.sectors for m in range(nf)], np.uint64)
        pd = _arr(pd, np.uint64)
        self.lam = torch.empty(self.size, dtype=torch.float64, device=dev)
        op.ctx.call("xr_sector_diag", nf, len(self.sectors), _ptr(dims), _ptr(offs), _ptr(e), _ptr(pd),
                    self.lam.data_ptr())

    def _pair_index(self, m, n, cm, cn):
        """Pair indices a * d_n + b of the states with charges (cm, cn)."""
        sm, nm = self.blocks_of[m][cm]
        sn, nn = self.blocks_of[n][cn]
        return ((np.arange(sm, sm + nm)[:, None]) * self.op.dims[n] + np.arange(sn, sn + nn)[None, :]).ravel()

    def _view(self, X, i):
        return X[int(self.offsets[i]):int(self.offsets[i + 1])]

    # -- launch plans ---------------------------------------------------------
    def _axis_plan(self, mats, trans, b, off):
        """Per fragment axis m: the arguments of xr_axis_apply_multi for every
        tuple (byte offsets of the tuple's block, the same in source and
        destination), plus the tuples whose axis is longer than 64."""
        nf = self.op.nf
        out = []
        for m in range(nf):
            fast = [i for i in range(len(self.sectors)) if self.shapes[i][m] <= 64]
            out.append({
                "U": _arr([mats[m][self.sectors[i][m]].data_ptr() for i in fast], np.uint64),
                "n": _arr([self.shapes[i][m] for i in fast], np.int32),
                "tr": _arr([trans] * len(fast), np.int32),
                "pre": _arr([int(np.prod(self.shapes[i][:m])) for i in fast], np.int64),
                "post": _arr([int(np.prod(self.shapes[i][m + 1:])) * b for i in fast], np.int64),
                "off": _arr([off[i] for i in fast], np.int64),
                "slow": [i for i in range(len(self.sectors)) if self.shapes[i][m] > 64]})
        return out

    def _pack_plan(self, items):
        """items: [(shape, ncol, f0, f1, nrows, full byte offset, packed byte
        offset, ld)] -> the arguments of xr_axes_pack_multi (offsets still
        relative to the full / packed bases)."""
        dims = np.zeros((len(items), 8), dtype=np.int64)
        for q, it in enumerate(items):
            dims[q, :len(it[0])] = it[0]
        return {"k": len(items), "nf": _arr([len(it[0]) for it in items], np.int32), "dims": dims.reshape(-1),
                "ncol": _arr([it[1] for it in items], np.int64), "f0": _arr([it[2] for it in items], np.int32),
                "f1": _arr([it[3] for it in items], np.int32), "nrows": _arr([it[4] for it in items], np.int64),
                "full": _arr([it[5] for it in items], np.int64), "packed": _arr([it[6] for it in items], np.int64),
                "ld": _arr([it[7] for it in items], np.int64)}

    def _plan(self, b):
        """Every launch argument of one application to b columns, built once
        per b: numpy arrays of shapes, fixed pointers (monomer matrices,
        square pair blocks) and byte offsets relative to the vectors and to
        the two scratch buffers (packed columns P, pair products T)."""
        pl = self._plans.get(b)
        if pl is not None:
            return pl
        off = [int(o) * b * 8 for o in self.offsets[:-1]]
        eigU = [{c: e[1] for c, e in self.eig[m].items()} for m in range(self.op.nf)]
        pl = {"mono": self._axis_plan(self.h1, 0, b, off), "fwd": self._axis_plan(eigU, 1, b, off),
              "inv": self._axis_plan(eigU, 0, b, off)}
        # pair terms, per pair (m, n) and charge sum s: the sector's tuples
        # with t_m + t_n = s are the row blocks (c_m, c_n) x column blocks
        # (charges of the other fragments, sorted) of one matrix P_s; the
        # product T_s = H_ss P_s has the same layout, so the unpack of every
        # tuple of the pair reads its own block of T_s
        nf = self.op.nf
        packs, grp, unpacks, pair_sums = [], {k: [] for k in ("M", "N", "K", "A", "lda", "B", "C")}, [], []
        poff = 0
        for (m, n) in self.op.pairs:
            rest = [a for a in range(nf) if a not in (m, n)]
            cols = {}
            for i, t in enumerate(self.sectors):
                cols.setdefault(t[m] + t[n], {})[tuple(t[a] for a in rest)] = i
            base, ups, sums = {}, [], []
            for s_, L in sorted(cols.items()):
                Hss, parts = self.sq[(m, n, s_)]
                W = 0
                for r in sorted(L):
                    sh = self.shapes[L[r]]
                    L[r] = W
                    W += int(np.prod([sh[a] for a in rest])) * b
                R = int(Hss.shape[0])
                rowoff = {(cm, cn): off for cm, cn, off, D in parts}
                base[s_] = (poff, W, rowoff, L)
                sums.append((s_, poff, R, W))
                for key, v in (("M", R), ("N", W), ("K", R), ("A", Hss.data_ptr()), ("lda", Hss.stride(0)),
                               ("B", poff), ("C", poff)):
                    grp[key].append(v)
                poff += 8 * R * W
            for i, t in enumerate(self.sectors):
                bo, W, rowoff, L = base[t[m] + t[n]]
                sh = self.shapes[i]
                at = bo + 8 * (rowoff[(t[m], t[n])] * W + L[tuple(t[a] for a in rest)])
                packs.append((sh, b, m, n, sh[m] * sh[n], off[i], at, W))
                ups.append((sh, b, m, n, sh[m] * sh[n], off[i], at, W))
            unpacks.append(self._pack_plan(ups))
            pair_sums.append(sums)
        pl["pack"] = self._pack_plan(packs)
        pl["pair_sums"] = pair_sums
        pl["grp"] = {k: _arr(v, np.uint64 if k == "A" else np.int64) for k, v in grp.items()}
        pl["unpacks"] = unpacks
        pl["P"] = pl["T"] = poff // 8
        self._plans[b] = pl
        return pl

    def _run_pack(self, pk, full, packed, mode):
        if pk["k"]:
            fp, pp = pk["full"] + full.data_ptr(), pk["packed"] + packed.data_ptr()   # alive through the call
            self.op.ctx.call("xr_axes_pack_multi", pk["k"], _ptr(pk["nf"]), _ptr(pk["dims"]), _ptr(pk["ncol"]),
                             _ptr(pk["f0"]), _ptr(pk["f1"]), None, _ptr(pk["nrows"]), _ptr(fp), _ptr(pp),
                             _ptr(pk["ld"]), int(mode), 1.0)

    def pair_density(self, x, m, n):
        """op.pair_density for a vector x of this sector: G_s = P_s P_s^T per
        charge sum s of the pair (P_s the packed tuples, as in matvec), then
        reordered from the square-block order into the charge-sorted order of
        the operator's pair blocks; sums without tuples here give zeros."""
        op = self.op
        pl = self._plan(1)
        pi = list(op.pairs).index((m, n))
        P = torch.empty(pl["P"], dtype=torch.float64, device=op.device)
        self._run_pack(pl["unpacks"][pi], x.reshape(-1, 1), P, 0)
        G = {}
        for s, boff, R, W in pl["pair_sums"][pi]:
            Ps = P[boff // 8: boff // 8 + R * W].view(R, W)
            G[s] = syrk(Ps)
        Pp = op.pairs[(m, n)]
        blocks = []
        for (start, cnt), q in zip(Pp["segs"], Pp["segq"]):
            if q not in G:
                blocks.append(torch.zeros(cnt, cnt, dtype=torch.float64, device=op.device))
                continue
            pos = getattr(self, "_segpos", {}).get((m, n, q))
            if pos is None:
                sq = op._sqidx[(m, n, q)]
                where = np.empty(op.dims[m] * op.dims[n], dtype=np.int64)
                where[sq] = np.arange(sq.size)
                pos = _i32(where[Pp["order_np"][start:start + cnt]], op.device)
                self.__dict__.setdefault("_segpos", {})[(m, n, q)] = pos
            blocks.append(G[q].index_select(0, pos).index_select(1, pos))
        return blocks

    def matvec(self, X, Y):
        """Y = H X for X, Y (size, b) in the sector layout."""
        op = self.op
        b = X.shape[1]
        pl = self._plan(b)
        for m in range(op.nf):
            self._axes_step(pl["mono"][m], self.h1, m, X, Y, b, 0, 0.0 if m == 0 else 1.0)
        # pair terms: every pair's tuples packed into its P_s matrices (one
        # launch set for all pairs), all products T_s = H_ss P_s in one
        # grouped GEMM, then one unpack launch set per pair (the tuples of one
        # pair are distinct targets; different pairs add into the same ones)
        P = torch.empty(pl["P"], dtype=torch.float64, device=op.device)
        T = torch.empty(pl["T"], dtype=torch.float64, device=op.device)
        self._run_pack(pl["pack"], X, P, 0)
        g = pl["grp"]
        k = len(g["M"])
        if k and op.grouped:
            Bp, Cp = g["B"] + P.data_ptr(), g["C"] + T.data_ptr()
            op.ctx.call("xr_gemm_grouped", k, 0, 0, _ptr(g["M"]), _ptr(g["N"]), _ptr(g["K"]), 1.0, _ptr(g["A"]),
                        _ptr(g["lda"]), _ptr(Bp), _ptr(g["N"]), _ptr(Cp), _ptr(g["N"]))
        elif k:
            for (m, n), sums in zip(op.pairs, pl["pair_sums"]):
                for s_, boff, R, W in sums:
                    o = boff // 8
                    mm(self.sq[(m, n, s_)][0], P[o:o + R * W].view(R, W), out=T[o:o + R * W].view(R, W))
        for up in pl["unpacks"]:
            self._run_pack(up, Y, T, 1)
        return Y

    def _axes_step(self, ap, mats, m, X, Y, b, trans, beta):
        """Applies mats[m][charge] (transposed when trans) along axis m of
        every tuple, X -> Y (+ beta Y)."""
        k = len(ap["n"])
        if k:
            src, dst = ap["off"] + X.data_ptr(), ap["off"] + Y.data_ptr()   # alive through the call
            self.op.ctx.call("xr_axis_apply_multi", k, _ptr(ap["U"]), _ptr(ap["n"]), _ptr(ap["tr"]),
                             _ptr(ap["pre"]), _ptr(ap["post"]), _ptr(src), _ptr(dst), float(beta))
        for i in ap["slow"]:
            sh = self.shapes[i]
            _axis_apply(mats[m][self.sectors[i][m]], self._view(X, i), self._view(Y, i), int(np.prod(sh[:m])),
                        sh[m], int(np.prod(sh[m + 1:])) * b, transA=trans, beta=beta)

    def _axes(self, X, inverse):
        """Applies U_m^T (or U_m when inverse) along every fragment axis of
        every tuple: the product eigenbasis of the monomer blocks."""
        b = X.shape[1]
        pl = self._plan(b)
        eigU = [{c: e[1] for c, e in self.eig[m].items()} for m in range(self.op.nf)]
        bufs = [torch.empty_like(X), torch.empty_like(X)]
        src = X                               # read only; the steps ping-pong between bufs
        for m in range(self.op.nf):
            ap = pl["inv" if inverse else "fwd"][m]
            dst = bufs[m % 2]
            self._axes_step(ap, eigU, m, src, dst, b, 0 if inverse else 1, 0.0)
            src = dst
        return src

    def precondition(self, R, theta):
        """(Hd - theta)^-1 R, Hd the diagonal of H in the product eigenbasis of
        the monomer blocks (monomer part exact, pair terms by their diagonal)."""
        Y = self._axes(R.contiguous(), False)
        return self._axes(precond_divide(Y, self.lam, theta, out=Y), True)

    def guess(self, b):
        """The b product eigenstates with the lowest diagonal energy."""
        b = min(b, self.size)
        # (a reduction for one vector; a device top-k sorts all entries)
        idx = torch.argmin(self.lam).view(1) if b == 1 else torch.topk(self.lam, b, largest=False).indices
        E = torch.zeros(self.size, b, dtype=torch.float64, device=self.op.device)
        E[idx, torch.arange(b, device=E.device)] = 1.0
        return self._axes(E, True)

    def to_product(self, x, out):
        """Writes the sector vector x into the product-layout vector out."""
        dims = self.op.dims
        full = out.view(dims)
        for i, t in enumerate(self.sectors):
            view = full
            for m in range(self.op.nf):
                st, n = self.blocks_of[m][t[m]]
                view = view.narrow(m, st, n)
            view.copy_(self._view(x, i).view(self.shapes[i]))
