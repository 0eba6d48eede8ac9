"""Autotuned GEMM configurations and the graph-captured XR block build.

* every tuned configuration gives the product of the untuned one (same
  algorithm, other tiling and K split: ~1e-15 relative);
* replaying the captured build after new coefficients are copied in equals
  an eager build with those coefficients, on several replays with eager
  work and synchronisation between them (graph memory nodes and library
  calls inside a capture have given wrong results on later replays).
"""
import ctypes

import numpy as np
import pytest
import torch

from qodeapplications_amd import synthetic as syn
from qodeapplications_amd._lib import context
from qodeapplications_amd.gemm import mm
from qodeapplications_amd.graph import CapturedBuild, autotune, gemm_shapes
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb,M,N,K", [(0, 0, 300, 200, 5000), (1, 0, 64, 96, 40000),
                                         (0, 1, 1500, 700, 300), (1, 1, 200, 1100, 900)])
def test_tuned_gemm_matches_heuristic(ta, tb, M, N, K):
    ctx = context(0)
    g = torch.Generator(device="cuda").manual_seed(M + N + K)
    A = torch.randn(*((K, M) if ta else (M, K)), dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(*((N, K) if tb else (K, N)), dtype=torch.float64, device="cuda", generator=g)
    ctx.call("xr_gemm_tune_reset")
    ref = mm(A, B, transA=ta, transB=tb)
    tile, splits, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    ctx.call("xr_gemm_tune", ta, tb, M, N, K, 1, 2, ctypes.byref(tile), ctypes.byref(splits), ctypes.byref(ms))
    # -1: the untuned dispatch was fastest and stays
    assert -1 <= tile.value <= 3 and (tile.value < 0 or splits.value >= 1) and ms.value > 0
    got = mm(A, B, transA=ta, transB=tb)
    scale = float(((A.abs().T if ta else A.abs()) @ (B.abs().T if tb else B.abs())).max())
    assert float((got - ref).abs().max()) < 1e-14 * scale
    ctx.call("xr_gemm_tune_reset")


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_grouped_gemm_matches_cpu(ta, tb):
    """xr_gemm_grouped over 30 products of uneven shapes (both tile grids,
    the single-product route, one empty product, edges that are not
    multiples of the tiles)."""
    ctx = context(0)
    rng = np.random.default_rng(ta * 2 + tb)
    shapes = [(int(rng.integers(1, 400)), int(rng.integers(1, 400)), int(rng.integers(1, 700)))
              for _ in range(29)] + [(0, 5, 7)]
    shapes[3] = (1024, 3072, 1024)
    shapes[4:7] = [(5000, 4, 300), (300, 40, 200), (50, 300, 100)]     # skinny, 128 x 64, small M
    probs = []
    for M, N, K in shapes:
        A = torch.randn(*((K, M) if ta else (M, K)), dtype=torch.float64)
        B = torch.randn(*((N, K) if tb else (K, N)), dtype=torch.float64)
        probs.append((A, B, torch.full((M, N), 7.0, dtype=torch.float64, device="cuda")))
    dev = [(A.cuda(), B.cuda(), C) for A, B, C in probs]
    k = len(dev)
    L, P = ctypes.c_int64, ctypes.c_void_p
    arr = lambda T, vals: (T * k)(*vals)
    ctx.call("xr_gemm_grouped", k, ta, tb, arr(L, [s[0] for s in shapes]), arr(L, [s[1] for s in shapes]),
             arr(L, [s[2] for s in shapes]), -0.5, arr(P, [A.data_ptr() for A, B, C in dev]),
             arr(L, [A.stride(0) for A, B, C in dev]), arr(P, [B.data_ptr() for A, B, C in dev]),
             arr(L, [B.stride(0) for A, B, C in dev]), arr(P, [C.data_ptr() for A, B, C in dev]),
             arr(L, [C.stride(0) for A, B, C in dev]))
    torch.cuda.synchronize()
    for (M, N, K), (A, B, _), (_, _, C) in zip(shapes, probs, dev):
        ref = -0.5 * ((A.T if ta else A) @ (B.T if tb else B))
        np.testing.assert_allclose(C.cpu().numpy(), ref.numpy(), atol=1e-12 * max(1, K) ** 0.5, rtol=1e-12)


def test_gemm_shapes_records_transposes():
    A = torch.randn(400, 300, dtype=torch.float64, device="cuda")
    B = torch.randn(500, 300, dtype=torch.float64, device="cuda")
    shapes = gemm_shapes(lambda: mm(A, B, transB=1))
    assert (0, 1, 400, 500, 300, 1) in shapes


def _coeff_keys(X):
    return [(m, c) for m, F in enumerate(X.frags) for c in F.charges]


def _new_coeffs(X, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = {}
    for m, c in _coeff_keys(X):
        z = X.frags[m].blocks[c].coeffs
        q, _ = torch.linalg.qr((z + 0.1 * torch.randn(z.shape, dtype=torch.float64, device="cuda",
                                                      generator=g)).T)
        out[(m, c)] = q.T.contiguous()
    return out


@pytest.mark.parametrize("cfg", ["c0", "six"])
def test_captured_build_replays_match_eager(cfg):
    if cfg == "c0":
        w = syn.config0(seed=0)
    else:   # six uneven small fragments, every pair block
        rng = np.random.default_rng(5)
        orbs = [2, 3, 2, 4, 3, 2]
        frags = [syn.make_fragment(n, [1, 2, 3], int(rng.integers(2, 4)), 2, rng) for n in orbs]
        w = syn.Workload("six", frags, syn.FragmentBlocked(orbs, syn.spatial_h(sum(orbs), rng),
                                                           syn.spatial_V(sum(orbs), rng)),
                         np.zeros((6, 6)), 5)
    X = XRSystem(w.fragments, w.integrals, w.nuc)
    cap = CapturedBuild(X, tune=(cfg == "c0"))
    ref_sys = XRSystem(w.fragments, w.integrals, w.nuc)
    for rep in range(3):
        z = _new_coeffs(X, 100 + rep)
        cap.set_coeffs(z)
        H1, H2, _ = cap.replay()
        torch.cuda.synchronize()
        # eager work and a synchronisation between replays
        for (m, c), t in z.items():
            ref_sys.frags[m].blocks[c] = ref_sys.frags[m].blocks[c].with_coeffs(t.clone())
        ref_sys.build_densities()
        R1, R2, _ = ref_sys.blocks()
        torch.cuda.synchronize()
        for a, b in zip(H1, R1):
            assert float((a - b).abs().max()) < 1e-12, rep
        nf = len(H1)
        for m in range(nf):
            for n in range(m + 1, nf):
                assert float((H2[m][n] - R2[m][n]).abs().max()) < 1e-12, (rep, m, n)
    context(0).call("xr_gemm_tune_reset")


def test_capture_pins_the_workspaces():
    """While a captured build exists, a call that would reallocate one of the
    context's workspaces fails with XR_EBUSY instead of freeing memory the
    graph replays; after release() the same call runs."""
    from qodeapplications_amd._lib import XR_EBUSY, XRError
    w = syn.config0(seed=0)
    X = XRSystem(w.fragments, w.integrals, w.nuc)
    context(0).call("xr_trim")        # workspaces sized by this build only
    cap = CapturedBuild(X, tune=False)
    with pytest.raises(XRError) as e:
        context(0).call("xr_trim")    # pinned
    assert e.value.code == XR_EBUSY
    # a product far larger than any of the build's: its staging needs a
    # larger workspace slot than the capture saw
    A = torch.randn(4096, 20000, dtype=torch.float64, device="cuda")
    B = torch.randn(20000, 4096, dtype=torch.float64, device="cuda")
    C = torch.zeros(4096, 4096, dtype=torch.float64, device="cuda")
    with pytest.raises(XRError) as e:
        mm(A, B, out=C, beta=1.0)
    assert e.value.code == XR_EBUSY
    H1, H2, _ = cap.replay()          # the graph still replays on its own memory
    torch.cuda.synchronize()
    cap.release()
    mm(A, B, out=C, beta=1.0)
    torch.cuda.synchronize()
    assert float((C - A @ B).abs().max()) < 1e-9 * float((A.abs() @ B.abs()).max())
    with pytest.raises(RuntimeError):
        cap.replay()
