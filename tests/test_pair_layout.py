"""Host logic of the pair-block build (no GPU): the fixed-charge block
layout used to scatter pair terms straight into the operator's blocks, and
the choice of which charge combinations are built."""
import numpy as np

from qodeapplications_amd.xr import XRSystem, sector_layout


def _segments(ch1, ch2):
    """The operator's ordering: pairs a * d2 + b stably sorted by charge sum."""
    q = (ch1[:, None] + ch2[None, :]).ravel()
    order = np.argsort(q, kind="stable")
    segs, segq, start = [], [], 0
    qs = q[order]
    for i in range(1, qs.size + 1):
        if i == qs.size or qs[i] != qs[start]:
            segs.append((start, i - start))
            segq.append(int(qs[start]))
            start = i
    return order, segs, segq


def test_sector_layout_matches_sorted_pairs():
    """Every pair's (block, row) from sector_layout equals its position in the
    stable charge-sum sort, for states of each charge scattered over the
    state order (not contiguous)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        ch1 = rng.integers(2, 6, size=rng.integers(3, 12))
        ch2 = rng.integers(2, 6, size=rng.integers(3, 12))
        order, segs, segq = _segments(ch1, ch2)
        off, rank2 = sector_layout(ch1, ch2, segq)
        d2 = ch2.size
        for (start, cnt), q in zip(segs, segq):
            for row, pair in enumerate(order[start:start + cnt]):
                a, b = divmod(int(pair), d2)
                assert ch1[a] + ch2[b] == q
                assert off[q][a] + rank2[b] == row


def test_combo_roles_pair_up():
    """Each charge combination that is not its own transpose is built exactly
    once (source) and its transpose skipped (derived); self-transposed ones
    keep both fragments' charges."""
    charges = [3, 4, 5]
    combos = [((a, b), (c, d)) for a in charges for b in charges for c in charges for d in charges
              if a + c == b + d and abs(a - b) <= 2]
    roles = {k: XRSystem.combo_role(*k) for k in combos}
    for (k1, k2), role in roles.items():
        kt = ((k1[1], k1[0]), (k2[1], k2[0]))
        assert kt in roles
        if role == "self":
            assert k1[0] == k1[1] and k2[0] == k2[1]
        else:
            assert {role, roles[kt]} == {"source", "derived"}
