"""Vector kernels of the block-Davidson solver (xr_precond_divide,
xr_residual, xr_column_norms, xr_transpose) against the torch expressions
they replace, and the fixed-order column sums reproducing bit for bit."""
import numpy as np
import pytest
import torch

from qodeapplications_amd._lib import XR_EUNSUPPORTED, context, ptr
from qodeapplications_amd.eigen import column_sumsq, precond_divide, residual, store_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,b", [(1, 1), (1000, 3), (100_003, 8), (70_000, 32)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_precond_divide_matches_torch(n, b, sign):
    g = torch.Generator().manual_seed(n + b)
    X = torch.randn(n, b, dtype=torch.float64, generator=g).cuda()
    lam = torch.randn(n, dtype=torch.float64, generator=g).cuda()
    theta = torch.randn(b, dtype=torch.float64, generator=g).cuda()
    lam[::7] = theta[0]                  # exact zeros of the denominator hit the floor
    if n > 1:
        lam[1] = theta[0] + 3e-9         # below the floor, positive
    d = sign * (lam[:, None] - theta[None, :])
    ref = X / torch.where(d.abs() < 1e-8, torch.full_like(d, 1e-8), d)
    out = precond_divide(X, lam, theta, sign=sign)
    Y = X.clone()
    precond_divide(Y, lam, theta, sign=sign, out=Y)          # in place
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    assert torch.equal(Y, ref)


@pytest.mark.parametrize("n,b,k", [(1, 1, 1), (5000, 4, 2), (300_001, 8, 8), (20_000, 32, 3), (1000, 3, 0)])
def test_residual_and_norms(n, b, k):
    g = torch.Generator().manual_seed(7 * n + b)
    X = torch.randn(n, b, dtype=torch.float64, generator=g).cuda()
    AX = torch.randn(n, b, dtype=torch.float64, generator=g).cuda()
    theta = torch.randn(b, dtype=torch.float64, generator=g).cuda()
    R, nr = residual(X, AX, theta, k)
    torch.cuda.synchronize()
    Rref = AX - X * theta
    # one rounding of X theta may be fused away (at most an ulp of either term)
    assert float((R - Rref).abs().max()) <= 2.3e-16 * float((X * theta).abs().max() + Rref.abs().max())
    if k:
        ref = torch.linalg.norm(Rref[:, :k], dim=0)
        assert float(((nr - ref).abs() / ref).max()) <= 1e-13
        R2, nr2 = residual(X, AX, theta, k)
        torch.cuda.synchronize()
        assert torch.equal(nr, nr2) and torch.equal(R, R2)      # fixed-order sums


def test_column_sumsq_and_strided_rows():
    g = torch.Generator().manual_seed(3)
    buf = torch.randn(50_000, 12, dtype=torch.float64, generator=g).cuda()
    A = buf[:, :7]                           # row stride 12
    s = column_sumsq(A)
    torch.cuda.synchronize()
    ref = (A * A).sum(0)
    assert float(((s - ref).abs() / ref).max()) <= 1e-13


@pytest.mark.parametrize("n,b", [(1, 1), (33, 5), (100_000, 8), (2_200_000, 2)])
def test_store_rows_is_the_transpose(n, b):
    g = torch.Generator().manual_seed(n)
    A = torch.randn(n, b, dtype=torch.float64, generator=g).cuda()
    Vt = torch.full((b + 3, n), 5.0, dtype=torch.float64, device="cuda")
    store_rows(Vt[1:1 + b], A)
    torch.cuda.synchronize()
    assert torch.equal(Vt[1:1 + b], A.T)
    assert bool((Vt[0] == 5.0).all()) and bool((Vt[1 + b:] == 5.0).all())


def test_solver_kernels_reject_wide_blocks():
    ctx = context(0)
    A = torch.zeros(10, 40, dtype=torch.float64, device="cuda")
    out = torch.zeros(40, dtype=torch.float64, device="cuda")
    rc = ctx.L.xr_column_norms(ctx.h, 10, 40, ptr(A), 40, ptr(out), 1)
    assert rc == XR_EUNSUPPORTED


@pytest.mark.parametrize("perm", [(0, 1, 3, 2), (1, 0, 2, 3), (0, 2, 1, 3), (3, 2, 1, 0), (0, 1, 2, 3)])
def test_permute4_matches_torch(perm):
    from qodeapplications_amd.gemm import permute4
    g = torch.Generator().manual_seed(sum(perm) + perm[0])
    src = torch.randn(5, 7, 3, 11, dtype=torch.float64, generator=g).cuda()
    out = permute4(src, perm)
    acc = torch.randn(out.shape, dtype=torch.float64, generator=g).cuda()
    acc0 = acc.clone()
    permute4(src, perm, out=acc, accumulate=True)
    torch.cuda.synchronize()
    ref = src.permute(*perm).contiguous()
    assert torch.equal(out, ref)
    assert torch.equal(acc, acc0 + ref)


@pytest.mark.parametrize("trans,n,K", [(0, 1, 1), (0, 130, 77), (1, 130, 77), (0, 3072, 3072), (1, 257, 4000)])
def test_syrk_matches_full_product(trans, n, K):
    """The symmetric product equals the full product to rounding and is
    exactly symmetric; rows of A given with a wider leading dimension."""
    from qodeapplications_amd.gemm import mm, syrk
    g = torch.Generator().manual_seed(n + K)
    shape = (K, n) if trans else (n, K)
    buf = torch.randn(shape[0], shape[1] + 3, dtype=torch.float64, generator=g).cuda()
    A = buf[:, :shape[1]]
    C = syrk(A, trans=trans, alpha=0.5)
    ref = mm(A, A, transA=1, alpha=0.5) if trans else mm(A, A, transB=1, alpha=0.5)
    torch.cuda.synchronize()
    assert torch.equal(C, C.T)
    scale = float(ref.abs().max())
    assert float((C - ref).abs().max()) <= 1e-13 * max(1.0, scale) * max(1, K) ** 0.5


def test_column_dots_and_sector_diag():
    """xr_column_dots against torch (row counts across chunk boundaries,
    leading dimensions above cols); xr_sector_diag against the broadcast sum
    of the monomer and pair diagonals for three fragments, two tuples."""
    ctx = context(0)
    g = torch.Generator().manual_seed(31)
    for rows, cols in [(1, 1), (64, 300), (65, 7), (1000, 1024)]:
        A = torch.randn(rows, cols + 2, dtype=torch.float64, generator=g).cuda()
        B = torch.randn(rows, cols + 5, dtype=torch.float64, generator=g).cuda()
        out = torch.empty(cols, dtype=torch.float64, device="cuda")
        ctx.call("xr_column_dots", rows, cols, ptr(A), cols + 2, ptr(B), cols + 5, ptr(out))
        ref = (A[:, :cols] * B[:, :cols]).sum(0)
        assert float((out - ref).abs().max()) <= 1e-13 * rows
    shapes = [(3, 4, 5), (2, 1, 6)]
    e = [[torch.randn(d, dtype=torch.float64, generator=g).cuda() for d in sh] for sh in shapes]
    pairs = [(0, 1), (0, 2), (1, 2)]
    pd = [[torch.randn(sh[m], sh[n], dtype=torch.float64, generator=g).cuda() for (m, n) in pairs]
          for sh in shapes]
    sizes = [int(np.prod(sh)) for sh in shapes]
    offs = np.array([0, sizes[0]], dtype=np.int64)
    dims = np.array(shapes, dtype=np.int64).reshape(-1)
    ep = np.array([t.data_ptr() for row in e for t in row], dtype=np.uint64)
    pp = np.array([t.data_ptr() for row in pd for t in row], dtype=np.uint64)
    out = torch.empty(sum(sizes), dtype=torch.float64, device="cuda")
    ctx.call("xr_sector_diag", 3, 2, dims.ctypes.data, offs.ctypes.data, ep.ctypes.data, pp.ctypes.data, ptr(out))
    for q, sh in enumerate(shapes):
        ref = torch.zeros(sh, dtype=torch.float64, device="cuda")
        for m in range(3):
            vs = [1, 1, 1]
            vs[m] = sh[m]
            ref = ref + e[q][m].view(vs)
        for p, (m, n) in enumerate(pairs):
            vs = [1, 1, 1]
            vs[m], vs[n] = sh[m], sh[n]
            ref = ref + pd[q][p].view(vs)
        assert torch.equal(out[int(offs[q]):int(offs[q]) + sizes[q]], ref.reshape(-1)), q
