"""Host logic of the certified charge-sector bound (no GPU): the Lagrangian
dual over the pair decomposition of the XR operator
(xr_operator.charge_sector_bound) against brute force.

Random charge-conserving monomer terms h_m and pair terms V_mn on 2-4
fragments are assembled into the full H on the product space, and each
total-charge sector's lowest eigenvalue comes from dense diagonalisation.
The bound built from the pair blocks' lowest eigenvalues, with the monomer
terms shared 1/(nf - 1) per pair and the charge multiplier's coefficient
s/(nf - 1) per pair block of charge sum s, must
  * never exceed the sector's lowest eigenvalue;
  * equal it for two fragments (the one pair term is all of H);
  * beat the multiplier-free bound (mu = 0) on some sector, which shows the
    multiplier does work;
  * stay below it for any other split of the monomer terms over the pairs
    (random shares summing to 1 per fragment);
  * equal the optimum of the linear programme it is the dual of (scipy's
    linprog: convex weights x_ps >= 0 of each pair's blocks with
    sum_s x_ps = 1 and sum_ps x_ps c_ps = Q, minimising sum_ps x_ps lam_ps),
    which pins the maximisation over mu apart from its breakpoint search.
Two plausible slips (the multiplier's coefficient s instead of s/(nf - 1),
or the -mu Q term dropped) give a "bound" above some sector's lowest
eigenvalue on the systems below, which shows that the brute-force check
sees them.
"""
import itertools

import numpy as np
import pytest

from qodeapplications_amd.xr_operator import charge_sector_bound


def _system(rng, nf):
    """Per fragment: state charges in {-1, 0, 1} (1-2 states each) and a
    symmetric charge-diagonal h; per pair: a symmetric V conserving the
    pair's charge sum."""
    ch, h = [], []
    for _ in range(nf):
        q = np.concatenate([[c] * int(rng.integers(1, 3)) for c in (-1, 0, 1)])
        A = rng.standard_normal((q.size, q.size))
        A = (A + A.T) * (q[:, None] == q[None, :])
        A += np.diag(0.5 * q * q)                 # charged states lie higher
        ch.append(q)
        h.append(A)
    V = {}
    for m, n in itertools.combinations(range(nf), 2):
        s = (ch[m][:, None] + ch[n][None, :]).ravel()
        B = 0.6 * rng.standard_normal((s.size, s.size))
        V[(m, n)] = (B + B.T) * (s[:, None] == s[None, :])
    return ch, h, V


def _exact(ch, h, V):
    """Lowest eigenvalue of the assembled H in every total-charge sector."""
    nf = len(ch)
    idx = np.array(list(itertools.product(*[range(q.size) for q in ch])))
    n = idx.shape[0]
    H = np.zeros((n, n))
    for m in range(nf):
        same = np.ones((n, n), bool)
        for k in range(nf):
            if k != m:
                same &= idx[:, None, k] == idx[None, :, k]
        H += same * h[m][idx[:, None, m], idx[None, :, m]]
    for (m, n_), Vmn in V.items():
        same = np.ones((n, n), bool)
        for k in range(nf):
            if k not in (m, n_):
                same &= idx[:, None, k] == idx[None, :, k]
        pi = idx[:, m] * ch[n_].size + idx[:, n_]
        H += same * Vmn[pi[:, None], pi[None, :]]
    tot = sum(ch[m][idx[:, m]] for m in range(nf))
    out = {}
    for Q in sorted(set(tot.tolist())):
        sel = np.where(tot == Q)[0]
        out[Q] = float(np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0])
    return out


def _pair_lams(ch, h, V, shares=None):
    """Lowest eigenvalue of h_mn = V_mn + w_m,mn h_m + w_n,mn h_n per charge
    sum (shares w; uniform 1/(nf - 1) by default)."""
    nf = len(ch)
    lam = {}
    for (m, n), Vmn in V.items():
        dn = ch[n].size
        wm = shares[(m, (m, n))] if shares else 1.0 / (nf - 1)
        wn = shares[(n, (m, n))] if shares else 1.0 / (nf - 1)
        A = Vmn + wm * np.kron(h[m], np.eye(dn)) + wn * np.kron(np.eye(ch[m].size), h[n])
        s = (ch[m][:, None] + ch[n][None, :]).ravel()
        for q in sorted(set(s.tolist())):
            sel = np.where(s == q)[0]
            lam[(m, n, q)] = float(np.linalg.eigvalsh(A[np.ix_(sel, sel)])[0])
    return lam


def _lines(ch, lam, Q, coef):
    """Per pair: (lam, coef(s)) of every charge sum s that the other
    fragments' charges complete to Q."""
    nf = len(ch)
    out = []
    for m, n in itertools.combinations(range(nf), 2):
        rest = {0}
        for f in range(nf):
            if f not in (m, n):
                rest = {t + int(x) for t in rest for x in set(ch[f].tolist())}
        S = sorted({int(a) + int(b) for a in ch[m] for b in ch[n]})
        out.append([(lam[(m, n, s)], coef(s)) for s in S if Q - s in rest])
    return out


def _lp(lines, Q):
    from scipy.optimize import linprog
    c = np.array([lam for L in lines for lam, _ in L])
    A_eq = np.zeros((len(lines) + 1, c.size))
    b_eq = np.ones(len(lines) + 1)
    j = 0
    for p, L in enumerate(lines):
        for _, coef in L:
            A_eq[p, j] = 1.0
            A_eq[-1, j] = coef
            j += 1
    b_eq[-1] = Q
    r = linprog(c, A_eq=A_eq, b_eq=b_eq, bounds=(0, None), method="highs")
    assert r.status == 0
    return float(r.fun)


@pytest.mark.parametrize("nf,seed", [(2, 0), (2, 1), (3, 2), (3, 3), (4, 4), (4, 5)])
def test_sector_bound_against_brute_force(nf, seed):
    rng = np.random.default_rng(seed)
    ch, h, V = _system(rng, nf)
    exact = _exact(ch, h, V)
    lam = _pair_lams(ch, h, V)
    coef = lambda s: s / (nf - 1)          # noqa: E731
    b = {Q: charge_sector_bound(_lines(ch, lam, Q, coef), Q) for Q in exact}
    for Q in exact:
        assert np.isfinite(b[Q])
        assert b[Q] <= exact[Q] + 1e-10, (Q, b[Q], exact[Q])
        if nf == 2:
            assert abs(b[Q] - exact[Q]) < 1e-10, (Q, b[Q], exact[Q])
        assert abs(b[Q] - _lp(_lines(ch, lam, Q, coef), Q)) < 1e-9 * max(1.0, abs(b[Q])), Q
    if nf > 2:
        b0 = {Q: charge_sector_bound(_lines(ch, lam, Q, lambda s: 0.0), Q) for Q in exact}
        assert any(b[Q] > b0[Q] + 1e-6 for Q in exact)


@pytest.mark.parametrize("nf,seed", [(3, 6), (4, 7)])
def test_sector_bound_with_any_monomer_shares(nf, seed):
    rng = np.random.default_rng(seed)
    ch, h, V = _system(rng, nf)
    exact = _exact(ch, h, V)
    for _ in range(3):
        shares = {}
        for m in range(nf):
            ps = [p for p in V if m in p]
            w = rng.dirichlet(np.ones(len(ps)))
            shares.update({(m, p): float(x) for p, x in zip(ps, w)})
        lam = _pair_lams(ch, h, V, shares)
        for Q in exact:
            b = charge_sector_bound(_lines(ch, lam, Q, lambda s: s / (nf - 1)), Q)
            assert b <= exact[Q] + 1e-10, (Q, b, exact[Q])


def test_planted_slips_exceed_the_exact_value():
    """Each slip overshoots some sector's lowest eigenvalue on one of the two
    systems (the wrong coefficient on the 3-fragment one, the dropped -mu Q
    on both)."""
    caught = {"coef": False, "no_q": False}
    for nf, seed in [(3, 2), (4, 5)]:
        rng = np.random.default_rng(seed)
        ch, h, V = _system(rng, nf)
        exact = _exact(ch, h, V)
        lam = _pair_lams(ch, h, V)
        for Q in exact:
            if charge_sector_bound(_lines(ch, lam, Q, float), Q) > exact[Q] + 1e-3:
                caught["coef"] = True
            if charge_sector_bound(_lines(ch, lam, Q, lambda s: s / (nf - 1)), 0) > exact[Q] + 1e-3:
                caught["no_q"] = True
    assert all(caught.values()), caught


def test_sector_bound_edge_cases():
    # one line per term: f is flat in mu (consistent slopes) and equals the sum
    assert charge_sector_bound([[(1.0, 0.5)], [(2.0, 0.5)]], 1) == 3.0
    # a term without a certified block makes the sector unbounded below
    assert charge_sector_bound([[(1.0, 0.0)], [(float("-inf"), 0.0), (2.0, 1.0)]], 0) == float("-inf")
    assert charge_sector_bound([], 0) == float("-inf")
    # the maximum lies where two lines of a term cross: f(mu) = min(mu, 1 - mu), at mu = 1/2
    assert abs(charge_sector_bound([[(0.0, 1.0), (1.0, -1.0)]], 0) - 0.5) < 1e-15
