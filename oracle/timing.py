"""CPU baseline: the oracle timed on the host.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg starts this module as
a child process (`python -m oracle.timing --workload c4 --procs P`), so its
worker processes inherit no GPU state; it prints one JSON line.  Nothing in
the product package imports it.

What it times (the oracle is the C/numpy restatement of the reference's CPU
path, oracle/xr_oracle.c and oracle/oracle.py):

* c0 measured, one process: the whole plain-FP64 dimer iteration of the
  reference's general-XRCC driver -- build_density_tensors for both
  fragments (every density string, the ccaa one contracted per state pair by
  monomer_2e, general-XRCC/build_density_tensors.py:57-160), every element
  of H1 and H2 through build_matrix_elements (general-XRCC/build_H.py:37-188,
  one call per element as test_H.py:49-138 drives it), braket_loops + fci
  (general-XRCC/hamiltonian.py:21-83, excitonic.py:56-82);
* c0 measured, P processes: the same work as a fork pool over
  (fragment, bra charge, ket charge, op) density jobs and over row chunks
  of the matrix elements, the reference's multiprocessing.Pool style
  (general-XRCC/test_H.py:131);
* a workload too large to run (c3, c4) is projected: rank-1 density
  updates per second and seconds per matrix-element call of each
  H_contractions family are sampled on the workload's own fragments, then
  multiplied by the workload's counts (updates of every density the
  reference builds, charge-conserving pair elements).  The same method's
  projection of c0 is reported beside c0's measured time as its check.
  The dense XR eigensolve is not projected for c4 (an 84,934,656-sided dense
  matrix) and the gradient is not included anywhere.
"""
import argparse
import json
import multiprocessing as mp
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

OPS = orc.OPS_BY_CHG_DIFF

_G = {}          # workload state inherited by the fork-pool workers


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def workload(name):
    from qodeapplications_amd import synthetic as syn
    return {"c0": lambda: syn.config0(seed=0), "c3": lambda: syn.config3(seed=3),
            "c4": lambda: syn.config4(seed=4)}[name]()


def paths_per_config(op, n_ket, dim):
    """Surviving (operator index tuple) paths per ket determinant: the
    reference's nested loops skip Pauli-violating indices."""
    occ, paths = n_ket, 1
    for o in reversed(op):
        if o == "a":
            paths *= occ
            occ -= 1
        else:
            paths *= dim - occ
            occ += 1
    return paths


# ---------------------------------------------------------------------------
# c0 measured
# ---------------------------------------------------------------------------

def _density_entry(m, bc, kc, op):
    w = _G["w"]
    f = w.fragments[m]
    zb = np.asarray(f.states[bc].coeffs, dtype=np.float64)
    zk = np.asarray(f.states[kc].coeffs, dtype=np.float64)
    R = orc.density(op, f.n_orbs, f.n_core, zb, zk, f.states[kc].configs)
    if op == "ccaa":
        V = np.ascontiguousarray(w.integrals.V[m, m, m, m])
        n_so = 2 * f.n_orbs
        return np.array([[orc.monomer_2e(n_so, R[i, j], V) for j in range(R.shape[1])]
                         for i in range(R.shape[0])])
    return R


def _dens_job(job):
    return job, _density_entry(*job)


def _rows_job(job):
    kind, M, N, r0, r1 = job
    w = _G["w"]
    bme = orc.build_matrix_elements(w.fragments, w.integrals, w.nuc)
    si = [f.state_indices for f in w.fragments]
    if kind == "mono":
        return job, np.array([[bme.monomer(M, I, J) for J in si[M]] for I in si[M][r0:r1]])
    basis = [(a, b) for a in si[M] for b in si[N]]
    return job, np.array([[bme.dimer((M, N), I, J) for J in basis] for I in basis[r0:r1]])


def _assemble_rho(w, results):
    for f in w.fragments:
        f.rho = {op: {} for ops in OPS.values() for op in ops}
    for (m, bc, kc, op), R in results:
        f = w.fragments[m]
        if op == "ccaa":
            f.rho[op][(bc, kc)] = [[float(R[i, j]) for j in range(R.shape[1])] for i in range(R.shape[0])]
        else:
            f.rho[op][(bc, kc)] = [[R[i, j] for j in range(R.shape[1])] for i in range(R.shape[0])]


def density_jobs(w):
    jobs = []
    for m, f in enumerate(w.fragments):
        for bc in f.states:
            for kc in f.states:
                for op in OPS.get(bc - kc, ()):
                    n = paths_per_config(op, f.states[kc].configs.shape[1], 2 * f.n_orbs)
                    cost = n * f.states[kc].configs.shape[0]
                    jobs.append((cost, (m, bc, kc, op)))
    jobs.sort(key=lambda t: -t[0])
    return [j for _, j in jobs]


def measure_c0(procs):
    w = workload("c0")
    _G["w"] = w
    out = {}
    # one process, the reference's order of work
    t0 = time.perf_counter()
    for m, f in enumerate(w.fragments):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, w.integrals.V[m, m, m, m], 0)
    t1 = time.perf_counter()
    H = orc.compute_H_blocks(w.fragments, w.integrals, w.nuc)
    t2 = time.perf_counter()
    E0, _ = orc.fci(H)
    t3 = time.perf_counter()
    out["single"] = {"densities_s": t1 - t0, "blocks_s": t2 - t1, "fci_s": t3 - t2, "total_s": t3 - t0,
                     "E0": float(E0)}
    # P processes: density jobs, then row chunks of the matrix elements
    w = workload("c0")
    _G["w"] = w
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_dens_job, density_jobs(w), chunksize=1)
    _assemble_rho(w, res)
    t1 = time.perf_counter()
    si = [f.state_indices for f in w.fragments]
    row_jobs = []
    for M in range(len(si)):
        row_jobs += [("mono", M, None, r, min(r + 8, len(si[M]))) for r in range(0, len(si[M]), 8)]
    nb = len(si[0]) * len(si[1])
    step = max(1, nb // (8 * procs))
    row_jobs += [("pair", 0, 1, r, min(r + step, nb)) for r in range(0, nb, step)]
    with ctx.Pool(procs) as pool:
        rows = pool.map(_rows_job, row_jobs, chunksize=1)
    H1 = [np.zeros((len(s), len(s))) for s in si]
    H2 = np.zeros((nb, nb))
    for (kind, M, N, r0, r1), blk in rows:
        if kind == "mono":
            H1[M][r0:r1] = blk
        else:
            H2[r0:r1] = blk
    t2 = time.perf_counter()
    E0p, _ = orc.fci((H1, [[None, H2], [None, None]], None))
    t3 = time.perf_counter()
    out["pool"] = {"processes": procs, "densities_s": t1 - t0, "blocks_s": t2 - t1, "fci_s": t3 - t2,
                   "total_s": t3 - t0, "E0": float(E0p)}
    return out


# ---------------------------------------------------------------------------
# projection from sampled rates
# ---------------------------------------------------------------------------

def _sample_density_rate(f, bc, kc, op, budget):
    """Rank-1 updates per second of the oracle density on growing slices of
    the ket determinants (the whole block when it fits the budget)."""
    sb, sk = f.states[bc], f.states[kc]
    dim = 2 * f.n_orbs
    nq, spent, rate, units = 16, 0.0, None, 0
    per_q = paths_per_config(op, sk.configs.shape[1], dim) * sb.coeffs.shape[0] * sk.coeffs.shape[0]
    while True:
        nq = min(nq, sk.configs.shape[0])
        zk = np.ascontiguousarray(sk.coeffs[:, :nq])
        t0 = time.perf_counter()
        orc.density(op, f.n_orbs, f.n_core, np.asarray(sb.coeffs), zk, sk.configs[:nq])
        dt = time.perf_counter() - t0
        spent += dt
        units = per_q * nq
        rate = units / max(dt, 1e-9)
        if nq == sk.configs.shape[0] or spent > budget or dt > 0.25 * budget:
            return rate, nq
        nq *= 2


def _time_call(fn, budget):
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        dt = time.perf_counter() - t0
        if dt > budget or n >= 200000:
            return dt / n


def project(w, budget):
    """Projected one-process seconds of the reference path (densities and
    matrix elements) for workload w, from rates sampled on w's fragments."""
    L = orc.lib()
    import ctypes
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    I64 = ctypes.c_int64
    frags = w.fragments
    # density rates per operator string, sampled on the largest fragment's
    # charge pair of that string with the most ket determinants
    rates, sample = {}, {}
    ops = sorted({op for ops in OPS.values() for op in ops})
    f0 = max(frags, key=lambda f: f.n_orbs)
    for op in ops:
        best = None
        for bc in f0.states:
            for kc in f0.states:
                if op in OPS.get(bc - kc, ()):
                    n = f0.states[kc].configs.shape[0]
                    if best is None or n > best[0]:
                        best = (n, bc, kc)
        rates[op], sample[op] = _sample_density_rate(f0, best[1], best[2], op, budget / (2 * len(ops)))
    t_dens = 0.0
    n_ccaa_pairs = {}
    for m, f in enumerate(frags):
        dim = 2 * f.n_orbs
        for bc in f.states:
            for kc in f.states:
                for op in OPS.get(bc - kc, ()):
                    units = (paths_per_config(op, f.states[kc].configs.shape[1], dim) * f.states[kc].configs.shape[0]
                             * f.states[bc].coeffs.shape[0] * f.states[kc].coeffs.shape[0])
                    t_dens += units / rates[op]
                    if op == "ccaa":
                        n_ccaa_pairs[dim] = n_ccaa_pairs.get(dim, 0) + \
                            f.states[bc].coeffs.shape[0] * f.states[kc].coeffs.shape[0]
    # matrix-element call times per family and spin-orbital dimensions
    rng = np.random.default_rng(0)
    per_call = budget / 40.0
    cache = {}

    def r(*shape):
        return np.ascontiguousarray(rng.standard_normal(shape))

    def cost(family, n1, n2):
        key = (family, n1, n2)
        if key not in cache:
            if family == "mono2e":
                R4, V = r(n1, n1, n1, n1), r(n1, n1, n1, n1)
                cache[key] = _time_call(lambda: L.orc_monomer_2e(I64(n1), P(R4), P(V)), per_call)
            elif family == "mono1e":
                R2, h = r(n1, n1), r(n1, n1)
                cache[key] = _time_call(lambda: L.orc_monomer_1e(I64(n1), P(R2), P(h)), per_call)
            elif family == "ExEx":
                A, B, V = r(n1, n1), r(n2, n2), r(n1, n2, n1, n2)
                cache[key] = _time_call(lambda: L.orc_dimer_ExEx(I64(n1), I64(n2), P(A), P(B), P(V)), per_call)
            elif family == "1min1pls":
                c1, cca1, a2, caa2 = r(n1), r(n1, n1, n1), r(n2), r(n2, n2, n2)
                h, V1, V2 = r(n1, n2), r(n1, n1, n1, n2), r(n1, n2, n2, n2)
                cache[key] = _time_call(lambda: (L.orc_dimer_1min1pls_1e(I64(n1), I64(n2), P(c1), P(a2), P(h)),
                                                 L.orc_dimer_1min1pls_2e(I64(n1), I64(n2), P(c1), P(cca1), P(a2),
                                                                          P(caa2), P(V1), P(V2))), per_call)
            else:      # 2min2pls
                cc1, aa2, V = r(n1, n1), r(n2, n2), r(n1, n1, n2, n2)
                cache[key] = _time_call(lambda: L.orc_dimer_2min2pls(I64(n1), I64(n2), P(cc1), P(aa2), P(V)),
                                        per_call)
        return cache[key]

    t_mono2e = sum(n * cost("mono2e", dim, 0) for dim, n in n_ccaa_pairs.items())
    t_blocks = 0.0
    n_elem = 0
    for m, f in enumerate(frags):
        d = sum(s.coeffs.shape[0] for s in f.states.values())
        t_blocks += d * d * cost("mono1e", 2 * f.n_orbs, 0)
    for a in range(len(frags)):
        for b in range(a + 1, len(frags)):
            fa, fb = frags[a], frags[b]
            na, nb = 2 * fa.n_orbs, 2 * fb.n_orbs
            for c1b in fa.states:
                for c1k in fa.states:
                    dd = c1b - c1k
                    if abs(dd) > 2:
                        continue
                    for c2b in fb.states:
                        c2k = c1b + c2b - c1k
                        if c2k not in fb.states:
                            continue
                        cnt = (fa.states[c1b].coeffs.shape[0] * fa.states[c1k].coeffs.shape[0]
                               * fb.states[c2b].coeffs.shape[0] * fb.states[c2k].coeffs.shape[0])
                        fam = {0: "ExEx", 1: "1min1pls", -1: "1min1pls", 2: "2min2pls", -2: "2min2pls"}[dd]
                        t_blocks += cnt * cost(fam, na, nb)
                        n_elem += cnt
    return {"densities_s": t_dens + t_mono2e, "blocks_s": t_blocks, "total_s": t_dens + t_mono2e + t_blocks,
            "charge_conserving_pair_elements": n_elem,
            "density_rates_updates_per_s": {k: float(v) for k, v in rates.items()},
            "density_sample_ket_determinants": sample,
            "element_call_s": {f"{k[0]}({k[1]},{k[2]})": v for k, v in cache.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c4", choices=["c0", "c3", "c4"])
    ap.add_argument("--procs", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--budget", type=float, default=12.0, help="seconds of sampling for a projection")
    ap.add_argument("--no-c0", action="store_true")
    args = ap.parse_args()
    out = {"nproc": os.cpu_count(), "cpu_model": cpu_model(), "procs": args.procs}
    if not args.no_c0:
        out["c0_measured"] = measure_c0(args.procs)
        out["c0_projected"] = project(workload("c0"), args.budget)
        m = out["c0_measured"]["single"]
        out["c0_projection_over_measured"] = (out["c0_projected"]["total_s"]
                                              / (m["densities_s"] + m["blocks_s"]))
    if args.workload != "c0":
        out["projected"] = project(workload(args.workload), args.budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
