"""Split-int8 GEMM (the Ozaki scheme on the int8 matrix cores) and the XR
path in that precision.

The scheme cuts each row of op(A) / column of op(B), scaled by a power of two
below its largest magnitude, into 7 signed 7-bit digits; the digit products
of weight 2^(-7 t), t <= 6, are summed exactly in int32.  What pins it:
  * integer operands with |x| < 64 are represented exactly by their first
    digit, so the product must equal the integer product bit for bit (this
    also pins the int8 MFMA operand and result lane layouts);
  * the elementwise bound |C - AB| <= 36 K 2^-49 max|a_row| max|b_col|
    (dropped digit pairs and the digit remainders), checked against an FP64
    CPU product with its own rounding allowance;
  * the XR energies and blocks of the plain dimer against the FP64 oracle.
"""
import os

import numpy as np
import pytest
import torch

from qodeapplications_amd import synthetic as syn
from qodeapplications_amd._lib import XR_EINVAL, context, ptr
from qodeapplications_amd.gemm import mm
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu

S = 7


def _bound(Am, Bm, alpha):
    """36 K 2^-49 |alpha| max|a_row| max|b_col| plus the FP64 reference's own
    rounding (K 2^-53 sum |a||b| at most)."""
    K = Am.shape[1]
    ra = Am.abs().amax(1, keepdim=True)
    cb = Bm.abs().amax(0, keepdim=True)
    dig = (S + 2) * 4 * K * 2.0 ** (-7 * S) * abs(alpha) * (ra @ cb)
    fp = K * 2.0 ** -52 * abs(alpha) * (Am.abs() @ Bm.abs())
    return dig + fp


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(77, 130, 200), (64, 64, 64), (130, 5, 1000), (1, 1, 1)])
def test_split_i8_exact_on_small_integers(ta, tb, M, N, K):
    g = torch.Generator().manual_seed(M * 7 + N + K)
    A = torch.randint(-63, 64, (K, M) if ta else (M, K), generator=g).double()
    B = torch.randint(-63, 64, (N, K) if tb else (K, N), generator=g).double()
    Am, Bm = (A.T if ta else A), (B.T if tb else B)
    ref = Am @ Bm                       # exact: integer sums below 2^53
    C = mm(A.cuda(), B.cuda(), transA=ta, transB=tb, split="i8")
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), ref)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K,wide", [(300, 260, 700, False), (64, 64, 40000, False),
                                        (1000, 1200, 333, False), (200, 300, 517, True),
                                        (2900, 2900, 640, False), (32, 40, 70000, True),
                                        (129, 65, 131073, False)])
def test_split_i8_error_bound(ta, tb, M, N, K, wide):
    """Elementwise bound above, with alpha and beta (the staged reduction),
    K beyond one int32 chunk (70000, 131073: split into chunks of <= 65536),
    and `wide` operands spread over 2^-40..2^40 with zero rows and columns."""
    g = torch.Generator().manual_seed(M + N + K + 1)
    A = torch.randn(*((K, M) if ta else (M, K)), dtype=torch.float64, generator=g)
    B = torch.randn(*((N, K) if tb else (K, N)), dtype=torch.float64, generator=g)
    if wide:
        A *= torch.exp2(torch.randint(-40, 41, A.shape, generator=g).double())
        B *= torch.exp2(torch.randint(-40, 41, B.shape, generator=g).double())
        A[..., 5] = 0.0
        B[3] = 0.0
    C0 = torch.randn(M, N, dtype=torch.float64, generator=g)
    Am, Bm = (A.T if ta else A), (B.T if tb else B)
    ref = 0.5 * (Am @ Bm) + 2.0 * C0
    C = C0.cuda()
    mm(A.cuda(), B.cuda(), transA=ta, transB=tb, alpha=0.5, out=C, beta=2.0, split="i8")
    err = (C.cpu() - ref).abs()
    tol = _bound(Am, Bm, 0.5) + 2.0 ** -51 * ref.abs()
    assert bool((err <= tol).all()), float((err / tol).max())
    if not wide:
        # N(0, 1) operands: far inside the bound, at FP64-product level
        typ = 1e-13 * (Am.abs() @ Bm.abs()) + 1e-15 * C0.abs()
        assert bool((err <= typ).all()), float((err / typ).max())


def test_split_i8_direct_output_and_leading_dimension():
    """beta = 0 and one K chunk: the kernel writes C directly (no staging),
    here into a view with a wider leading dimension; the columns outside the
    view stay untouched."""
    g = torch.Generator().manual_seed(5)
    A = torch.randn(3000, 900, dtype=torch.float64, generator=g)
    B = torch.randn(900, 2500, dtype=torch.float64, generator=g)
    buf = torch.full((3000, 2600), 7.0, dtype=torch.float64, device="cuda")
    out = buf[:, :2500]
    mm(A.cuda(), B.cuda(), alpha=-1.5, out=out, split="i8")
    torch.cuda.synchronize()
    ref = -1.5 * (A @ B)
    err = (out.cpu() - ref).abs()
    assert bool((err <= _bound(A, B, 1.5) + 2.0 ** -51 * ref.abs()).all())
    assert bool((buf[:, 2500:] == 7.0).all())


def test_split_i8_tall_operand():
    """More 64-row blocks (65625) than a grid's y dimension holds: the digit
    passes walk a one-dimensional block list."""
    g = torch.Generator().manual_seed(13)
    M, N, K = 4_200_000, 8, 16
    A = torch.randn(M, K, dtype=torch.float64, generator=g)
    B = torch.randn(K, N, dtype=torch.float64, generator=g)
    C = mm(A.cuda(), B.cuda(), split="i8")
    torch.cuda.synchronize()
    ref = A @ B
    err = (C.cpu() - ref).abs()
    assert bool((err <= _bound(A, B, 1.0) + 2.0 ** -51 * ref.abs()).all())


def test_split_i8_batched_strides():
    ctx = context(0)
    g = torch.Generator().manual_seed(11)
    nb, M, N, K = 3, 150, 90, 333
    A = torch.randn(nb, M, K, dtype=torch.float64, generator=g)
    B = torch.randn(nb, N, K, dtype=torch.float64, generator=g)
    C = torch.randn(nb, M, N, dtype=torch.float64, generator=g)
    ref = 2.0 * (A @ B.transpose(1, 2)) - C
    Ad, Bd, Cd = A.cuda(), B.cuda(), C.cuda()
    ctx.call("xr_gemm_split_i8", 0, 1, M, N, K, 2.0, ptr(Ad), K, M * K, ptr(Bd), K, N * K, -1.0, ptr(Cd), N,
             M * N, nb)
    torch.cuda.synchronize()
    for b in range(nb):
        err = (Cd[b].cpu() - ref[b]).abs()
        assert bool((err <= _bound(A[b], B[b].T, 2.0) + 2.0 ** -51 * ref[b].abs()).all()), b


def test_mode_routing():
    """xr_gemm_mode in split-int8 mode: a small product (below 2^33 flops)
    runs the FP64 kernel (bit-identical to xr_gemm_f64), a large one the
    int8 kernel (bit-identical to xr_gemm_split_i8); split-FP32 mode is the
    split-FP32 product."""
    g = torch.Generator().manual_seed(17)
    for (M, N, K) in [(300, 260, 700), (2048, 2048, 1100)]:
        A = torch.randn(M, K, dtype=torch.float64, generator=g).cuda()
        B = torch.randn(K, N, dtype=torch.float64, generator=g).cuda()
        routed = mm(A, B, mode="split_i8")
        f64 = mm(A, B)
        i8 = mm(A, B, split="i8")
        f32 = mm(A, B, split="f32")
        torch.cuda.synchronize()
        big = 2.0 * M * N * K >= 2.0 ** 33
        assert torch.equal(routed, i8 if big else f64), (M, N, K)
        assert not torch.equal(i8, f64)
        assert torch.equal(mm(A, B, mode="split_f32"), f32)
        assert torch.equal(mm(A, B, mode="f64"), f64)


def test_split_i8_rejects_bad_sizes():
    ctx = context(0)
    L = ctx.L
    A = torch.zeros(4, 4, dtype=torch.float64, device="cuda")
    rc = L.xr_gemm_split_i8(ctx.h, 0, 0, -1, 4, 4, 1.0, ptr(A), 4, 0, ptr(A), 4, 0, 0.0, ptr(A), 4, 0, 1)
    assert rc == XR_EINVAL
    assert b"negative" in L.xr_last_error(ctx.h)


def test_split_i8_xr_energy():
    """Split-int8 mode on the plain dimer (9+9 spatial orbitals, 3/4/5
    electrons, 8 states per block): monomer and pair blocks within 1e-10 of
    the FP64 oracle, the lowest 10 XR eigenvalues within 1e-10 Eh.  At this
    size most two-electron products are memory bound and the mode's shape
    rule keeps them on FP64; the int8 products of the XR path are pinned at
    the 2 x 18-orbital dimer (test_gpu_c3.py)."""
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "c0_oracle.npz"))
    w = syn.config0(seed=0)
    X = XRSystem(w.fragments, w.integrals, w.nuc, precision="split_i8")
    H = X.blocks()
    for m in range(2):
        assert float(np.abs(H[0][m].cpu().numpy() - ref[f"H1_{m}"]).max()) < 1e-10
    assert float(np.abs(H[1][0][1].cpu().numpy() - ref["H2"]).max()) < 1e-10
    M = X.xr_matrix(H=H)
    assert float((M - M.T).abs().max()) == 0.0
    ev = torch.linalg.eigvalsh(M)[:10].cpu().numpy()
    assert float(np.abs(ev - ref["evals"]).max()) < 1e-10


def test_split_i8_pair_program_matches_per_combination():
    w = syn.config0(seed=0)
    X = XRSystem(w.fragments, w.integrals, w.nuc, precision="split_i8")
    X.build_densities()
    H = X.pair_matrices([(0, 1)])[(0, 1)]
    ref = X.pair_matrix(0, 1)
    Xd = XRSystem(w.fragments, w.integrals, w.nuc)
    Xd.build_densities()
    exact = Xd.pair_matrix(0, 1)
    torch.cuda.synchronize()
    scale = float(exact.abs().max())
    assert float((H - ref).abs().max()) <= 1e-12 * scale
    assert float((H - exact).abs().max()) <= 1e-11 * scale
