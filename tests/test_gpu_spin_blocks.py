"""Spin-orbital integral blocks from the spatial integrals (xr_spin_blocks)
and their adjoint (xr_spin_blocks_adjoint).

Forward: equal to the numpy construction of the synthetic workloads
(synthetic.FragmentBlocked, written independently from the same definition,
V_pqrs = (1/4)(<pq|rs> - <pq|sr>), alpha spin orbitals first).  Adjoint: equal
to torch autograd through a plain torch restatement of the blocks (slices,
spin cases, concatenation) on random block gradients."""
import itertools

import numpy as np
import pytest
import torch

from qodeapplications_amd import synthetic as syn
from qodeapplications_amd.orbitals import OrbitalRotation, _SpinBlocks, _spin_desc, integral_keys

pytestmark = pytest.mark.gpu


def _system(orbs, seed):
    rng = np.random.default_rng(seed)
    n = sum(orbs)
    h = rng.standard_normal((n, n))
    h = h + h.T
    V = rng.standard_normal((n, n, n, n))
    V = V + V.transpose(1, 0, 2, 3)
    V = V + V.transpose(0, 1, 3, 2)
    V = V + V.transpose(2, 3, 0, 1)
    U = rng.standard_normal((len(orbs), n, n))
    U = U + U.transpose(0, 2, 1)
    return h, V, U


def _torch_blocks(h, V, U, keys, orbs):
    """Plain torch restatement (for the autograd reference)."""
    off = np.concatenate([[0], np.cumsum(orbs)]).astype(int)
    sl = lambda m: slice(int(off[m]), int(off[m + 1]))
    out = []
    for key in keys:
        if key[0] in ("T", "U"):
            A = h if key[0] == "T" else U[key[1]]
            m1, m2 = key[-2:]
            blk = A[sl(m1), sl(m2)]
            z = torch.zeros_like(blk)
            out.append(torch.cat([torch.cat([blk, z], 1), torch.cat([z, blk], 1)], 0))
        else:
            ms = key[1:]
            s = [sl(m) for m in ms]
            C = V[s[0], s[2], s[1], s[3]].permute(0, 2, 1, 3)
            X = V[s[0], s[3], s[1], s[2]].permute(0, 2, 3, 1)
            cube = {}
            for sp in itertools.product((0, 1), repeat=4):
                blk = torch.zeros_like(C)
                if sp[0] == sp[2] and sp[1] == sp[3]:
                    blk = blk + C
                if sp[0] == sp[3] and sp[1] == sp[2]:
                    blk = blk - X
                cube[sp] = 0.25 * blk

            def cat(level, prefix):
                if level == 4:
                    return cube[prefix]
                return torch.cat([cat(level + 1, prefix + (b,)) for b in (0, 1)], dim=level)
            out.append(cat(0, ()))
    return out


@pytest.mark.parametrize("orbs", [(2, 3), (2, 3, 4), (3, 1, 2, 2)])
def test_spin_blocks_match_numpy_construction(orbs):
    h, V, U = _system(orbs, sum(orbs))
    keys = integral_keys(len(orbs))
    rot = OrbitalRotation(orbs, h, V, U)
    blocks = rot.blocks(rot.zero_kappas(), keys, at_zero=True)
    ref = syn.FragmentBlocked(orbs, h, V, U)
    torch.cuda.synchronize()
    for key in keys:
        got = blocks[key].detach().cpu().numpy()
        want = ref.block(key[0], key[1:])
        assert got.shape == want.shape, key
        assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max()), key


@pytest.mark.parametrize("orbs", [(2, 3), (3, 1, 2, 2)])
def test_spin_blocks_adjoint_matches_autograd(orbs):
    h, V, U = _system(orbs, 10 + sum(orbs))
    keys = integral_keys(len(orbs))
    dev = torch.device("cuda", 0)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True)
    h1, V1, U1 = t(h), t(V), t(U)
    h2, V2, U2 = t(h), t(V), t(U)
    off = np.concatenate([[0], np.cumsum(orbs)]).astype(int)
    lib = _SpinBlocks.apply(_spin_desc(keys, list(orbs), off), h1, V1, U1)
    ref = _torch_blocks(h2, V2, U2, keys, orbs)
    g = torch.Generator().manual_seed(5)
    grads = [torch.randn(r.shape, dtype=torch.float64, generator=g).to(dev) for r in ref]
    grads[3] = None                          # a block without a gradient
    torch.autograd.backward([x for x, gg in zip(lib, grads) if gg is not None], [gg for gg in grads if gg is not None])
    torch.autograd.backward([x for x, gg in zip(ref, grads) if gg is not None], [gg for gg in grads if gg is not None])
    torch.cuda.synchronize()
    for a, b in ((h1, h2), (V1, V2), (U1, U2)):
        assert float((a.grad - b.grad).abs().max()) <= 1e-13 * max(1.0, float(b.grad.abs().max()))
