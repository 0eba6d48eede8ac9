"""Certified lower bounds of the charge sectors' lowest XR eigenvalues
(xr_pair_dress, xr_lanczos, xr_add_diagonal and XROperator.sector_bounds).

  * xr_pair_dress equals the same sums formed with torch;
  * xr_lanczos: on matrices of known spectrum the lowest eigenvalue of the
    Lanczos tridiagonal is the lowest eigenvalue once the Krylov space is
    the whole space (6 x 6 in 6 steps), and converges to it from above in
    40 steps on 700 x 700 and 3000 x 3000 matrices with a gap at the bottom
    of the spectrum; anorm is the largest absolute row sum;
  * xr_add_diagonal shifts exactly the diagonals;
  * xr_cholesky_test: on a batch of sizes 1..2000 (row strides above n
    included) the outcome for A - t I at t = lambda_min -/+ 1e-3 is positive
    definite / not, as LAPACK's dpotrf says, and a diagonally dominant matrix
    with one negative diagonal entry at column k fails exactly at k + 1
    (in the first, second and last panel); a NaN fails;
  * sector_bounds on 3- and 4-fragment chains: every certified block bound
    lies below the block's exact lowest eigenvalue (torch eigvalsh) and
    within 0.05 of it, and every sector's bound below the sector's exact
    lowest eigenvalue from the dense operator;
  * lowest() returns the same root with and without the bounds.
"""
import itertools

import numpy as np
import pytest
import torch

from qodeapplications_amd import synthetic as syn
from qodeapplications_amd._lib import XR_EINVAL, context, ptr
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu


def _chain(seed, orbs, ns):
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, [1, 2, 3], ns, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    U = np.stack([0.2 * syn.spatial_h(n_sp, rng) for _ in orbs])
    nf = len(orbs)
    nuc = np.array([[0.0 if m == n else 0.3 / abs(m - n) for n in range(nf)] for m in range(nf)])
    return XRSystem(frags, syn.FragmentBlocked(list(orbs), h, V, U), nuc)


def test_pair_dress_matches_torch():
    ctx = context(0)
    g = torch.Generator().manual_seed(3)
    dm, dn = 7, 9
    idx = torch.randperm(dm * dn, generator=g)[:50]
    pa, pb = (idx // dn).int().cuda(), (idx % dn).int().cuda()
    cnt = idx.numel()
    B = torch.randn(cnt, cnt + 5, dtype=torch.float64, generator=g).cuda()
    hm = torch.randn(dm, dm, dtype=torch.float64, generator=g).cuda()
    hn = torch.randn(dn, dn, dtype=torch.float64, generator=g).cuda()
    out = torch.full((cnt, cnt + 3), 9.0, dtype=torch.float64, device="cuda")
    wm, wn = 0.37, 0.81
    ctx.call("xr_pair_dress", cnt, ptr(B), cnt + 5, ptr(pa), ptr(pb), ptr(hm), dm, ptr(hn), dn, wm, wn, ptr(out),
             cnt + 3)
    Bs = B[:, :cnt]
    a, b = pa.long(), pb.long()
    hms, hns = 0.5 * (hm + hm.T), 0.5 * (hn + hn.T)
    ref = 0.5 * (Bs + Bs.T) + wm * hms[a][:, a] * (b[:, None] == b[None, :]) \
        + wn * hns[b][:, b] * (a[:, None] == a[None, :])
    assert float((out[:, :cnt] - ref).abs().max()) < 1e-14
    assert bool((out[:, cnt:] == 9.0).all())


def _known(n, low, gen):
    """Q diag(spectrum) Q^T with the lowest eigenvalues `low` well below the rest."""
    Q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, generator=gen))
    ev = torch.cat([torch.tensor(low, dtype=torch.float64),
                    torch.linspace(0.0, 2.0, n - len(low), dtype=torch.float64)])
    return (Q * ev) @ Q.T, float(min(low))


def _ritz_low(a, b):
    k = len(a)
    T = np.diag(a) + np.diag(b[:k - 1], 1) + np.diag(b[:k - 1], -1)
    return float(np.linalg.eigvalsh(T)[0])


def test_lanczos_lowest_ritz_values():
    ctx = context(0)
    gen = torch.Generator().manual_seed(8)
    mats = [_known(6, [-1.5], gen), _known(700, [-1.0, -0.8], gen), _known(3000, [-0.7], gen)]
    steps = 40
    As = [M.cuda() for M, _ in mats]
    ns = np.array([A.shape[0] for A in As], dtype=np.int64)
    ps = np.array([A.data_ptr() for A in As], dtype=np.uint64)
    nb = len(As)
    alpha = torch.empty(nb, steps, dtype=torch.float64, device="cuda")
    beta = torch.empty_like(alpha)
    anorm = torch.empty(nb, dtype=torch.float64, device="cuda")
    ctx.call("xr_lanczos", nb, ns.ctypes.data, ps.ctypes.data, ns.ctypes.data, steps, ptr(alpha), ptr(beta),
             ptr(anorm))
    a, b, an = alpha.cpu().numpy(), beta.cpu().numpy(), anorm.cpu().numpy()
    for j, (M, lo) in enumerate(mats):
        assert abs(an[j] - float(M.abs().sum(1).max())) <= 1e-13 * an[j]
        k = steps if j else 6                   # the 6 x 6 Krylov space is complete after 6 steps
        th = _ritz_low(a[j, :k], b[j, :k])
        assert th >= lo - 1e-10
        assert th - lo < (1e-10 if j == 0 else 1e-6), (j, th, lo)


def test_lanczos_and_diagonal_reject_bad_arguments():
    ctx = context(0)
    L = ctx.L
    n = np.array([4], dtype=np.int64)
    p = np.array([0], dtype=np.uint64)
    assert L.xr_lanczos(ctx.h, 65, n.ctypes.data, p.ctypes.data, n.ctypes.data, 5, 0, 0, 0) == XR_EINVAL
    assert L.xr_lanczos(ctx.h, 1, n.ctypes.data, p.ctypes.data, n.ctypes.data, 0, 0, 0, 0) == XR_EINVAL
    assert L.xr_add_diagonal(ctx.h, 65, n.ctypes.data, p.ctypes.data, n.ctypes.data, 0) == XR_EINVAL


def test_add_diagonal_shifts_only_the_diagonals():
    ctx = context(0)
    A = torch.randn(5, 7, dtype=torch.float64).cuda()          # 5 x 5 view with row stride 7
    B = torch.randn(3, 3, dtype=torch.float64).cuda()
    A0, B0 = A.clone(), B.clone()
    ns = np.array([5, 3], dtype=np.int64)
    lds = np.array([7, 3], dtype=np.int64)
    ps = np.array([A.data_ptr(), B.data_ptr()], dtype=np.uint64)
    sh = np.array([0.25, -2.0])
    ctx.call("xr_add_diagonal", 2, ns.ctypes.data, ps.ctypes.data, lds.ctypes.data, sh.ctypes.data)
    refA, refB = A0.clone(), B0.clone()
    refA[:, :5] += 0.25 * torch.eye(5, dtype=torch.float64, device="cuda")
    refB += -2.0 * torch.eye(3, dtype=torch.float64, device="cuda")
    assert torch.equal(A, refA) and torch.equal(B, refB)


def _chol(mats, lds=None):
    ctx = context(0)
    lds = lds or [A.shape[0] for A in mats]
    bufs = []
    for A, ld in zip(mats, lds):
        n = A.shape[0]
        B = torch.full((n, ld), 7.0, dtype=torch.float64)
        B[:, :n] = A
        bufs.append(B.cuda())
    ns = np.array([A.shape[0] for A in mats], dtype=np.int64)
    lda = np.array(lds, dtype=np.int64)
    ps = np.array([B.data_ptr() for B in bufs], dtype=np.uint64)
    info = torch.full((len(mats),), -5, dtype=torch.int32, device="cuda")
    ctx.call("xr_cholesky_test", len(mats), ns.ctypes.data, ps.ctypes.data, lda.ctypes.data, ptr(info))
    return info.cpu().numpy()


def test_cholesky_test_outcomes():
    from scipy.linalg import lapack
    gen = torch.Generator().manual_seed(21)
    sizes = [1, 5, 64, 65, 130, 700, 2000]
    base = []
    for n in sizes:
        G = torch.randn(n, n, dtype=torch.float64, generator=gen)
        base.append(G @ G.T / n + 0.1 * torch.eye(n, dtype=torch.float64))
    lows = [float(torch.linalg.eigvalsh(A)[0]) for A in base]
    for sign in (-1.0, 1.0):
        mats = [A - (lo + sign * 1e-3) * torch.eye(A.shape[0], dtype=torch.float64) for A, lo in zip(base, lows)]
        info = _chol(mats, [A.shape[0] + (3 if j % 2 else 0) for j, A in enumerate(mats)])
        for j, A in enumerate(mats):
            ref = lapack.dpotrf(A.numpy(), lower=1)[1]
            assert (info[j] == 0) == (ref == 0) == (sign < 0), (sizes[j], sign, info[j], ref)
    mats = []
    for n, k in [(200, 10), (200, 100), (300, 299), (64, 63)]:
        A = 0.01 * torch.rand(n, n, dtype=torch.float64, generator=gen)
        A = A + A.T + 2.0 * torch.eye(n, dtype=torch.float64)
        A[k, k] = -1.0
        mats.append(A)
    N = base[4].clone()
    N[70, 3] = N[3, 70] = float("nan")
    mats.append(N)
    info = _chol(mats)
    assert list(info[:4]) == [11, 101, 300, 64]
    assert info[4] != 0


def _sector_lows(op):
    """Exact lowest eigenvalue per total charge from the dense operator."""
    M = op.dense()
    M = 0.5 * (M + M.T)
    digits = np.array(list(itertools.product(*[range(d) for d in op.dims])))
    tot = sum(op.charges[m][digits[:, m]] for m in range(op.nf))
    out = {}
    for Q in sorted(set(tot.tolist())):
        sel = torch.as_tensor(np.where(tot == Q)[0], device="cuda")
        out[Q] = float(torch.linalg.eigvalsh(M[sel][:, sel])[0])
    return out


@pytest.mark.parametrize("orbs,ns", [((3, 3, 3), 4), ((2, 3, 2, 3), 3)])
def test_sector_bounds_below_exact(orbs, ns):
    X = _chain(51, orbs, ns)
    X.build_densities()
    op = X.operator()
    bounds = op.sector_bounds()
    lows = _sector_lows(op)
    assert set(bounds) == set(lows)
    shares = op.monomer_shares
    assert op.block_bounds
    for (m, n, s), lb in op.block_bounds.items():
        P = op.pairs[(m, n)]
        i = P["segq"].index(s)
        start, cnt = P["segs"][i]
        idx = torch.as_tensor(P["order_np"][start:start + cnt], device="cuda")
        a, b = idx // op.dims[n], idx % op.dims[n]
        B = P["blocks"][i]
        hm, hn = 0.5 * (op.H1[m] + op.H1[m].T), 0.5 * (op.H1[n] + op.H1[n].T)
        A = 0.5 * (B + B.T) + shares[(m, (m, n))] * hm[a][:, a] * (b[:, None] == b[None, :]) \
            + shares[(n, (m, n))] * hn[b][:, b] * (a[:, None] == a[None, :])
        lam = float(torch.linalg.eigvalsh(A)[0])
        assert lb <= lam + 1e-12, (m, n, s, lb, lam)
        assert lam - lb < 0.05, (m, n, s, lb, lam)
    for Q, lb in bounds.items():
        print(f"chain {orbs}: Q={Q:+d} bound {lb:.6f} lowest {lows[Q]:.6f}", flush=True)
        assert lb <= lows[Q] + 1e-10, (Q, lb, lows[Q])
        assert np.isfinite(lb)


def test_lowest_same_root_with_bounds():
    X = _chain(42, (4, 4, 4), 7)
    X.build_densities()
    op = X.operator()
    op.log = []
    w1, _ = op.lowest(k=1)
    log1 = list(op.log)
    op.log = []
    w0, _ = op.lowest(k=1, bound=False)
    for row in log1:
        print("bounded solve: Q=%+d size=%d values=%s residual=%.2e iters=%d %s" % row, flush=True)
    assert abs(float(w1[0]) - float(w0[0])) <= 1e-12 * max(1.0, abs(float(w0[0])))
    for row in log1:
        if row[5] == "bounded-above":
            assert row[2][0] > float(w1[0])


def test_cholesky_test_rejects_bad_arguments():
    ctx = context(0)
    L = ctx.L
    A = torch.eye(4, dtype=torch.float64, device="cuda")
    n = np.array([4], dtype=np.int64)
    p = np.array([A.data_ptr()], dtype=np.uint64)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert L.xr_cholesky_test(ctx.h, 65, n.ctypes.data, p.ctypes.data, n.ctypes.data, ptr(info)) == XR_EINVAL
    small = np.array([3], dtype=np.int64)
    from qodeapplications_amd._lib import XR_ESHAPE
    assert L.xr_cholesky_test(ctx.h, 1, n.ctypes.data, p.ctypes.data, small.ctypes.data, ptr(info)) == XR_ESHAPE
    zero = np.array([0], dtype=np.int64)
    assert L.xr_cholesky_test(ctx.h, 1, zero.ctypes.data, p.ctypes.data, n.ctypes.data, ptr(info)) == XR_EINVAL


def test_no_bounds_without_every_pair_term():
    """An operator whose pair terms do not cover every pair gives no sector
    bounds (the decomposition needs each fragment in all its pairs) and its
    lowest root is the one without bounds."""
    from qodeapplications_amd.xr_operator import XROperator
    X = _chain(42, (3, 3, 3), 5)
    H1, H2, _ = X.blocks()
    pairs = {(0, 1): H2[0][1], (1, 2): H2[1][2]}          # (0, 2) left out
    op = XROperator(H1, pairs, X.state_charges())
    assert op.sector_bounds() == {}
    op.log = []
    w1, _ = op.lowest(k=1)
    assert op.bounds == {}
    assert all(r[5] != "bounded-above" for r in op.log)
    M = op.dense()
    ref = float(torch.linalg.eigvalsh(0.5 * (M + M.T))[0])
    assert abs(float(w1[0]) - ref) < 1e-9
