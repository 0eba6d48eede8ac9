"""XR H-build + gradient benchmark on one MI355X.

One step = one XR iteration on the split-precision dimer workload (two
fragments of 18 spatial orbitals, 3/4/5 electrons, 64 states per charge
block, seeded synthetic states and integrals): build every transition
density the XR blocks need, the monomer two-electron sigma terms, all
monomer and pair blocks, assemble the 36864 x 36864 XR matrix, solve for the
lowest eigenpair, and run the adjoint for dE/dz and dE/dkappa.  Inputs are
resident in HBM before the timed region.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_F64_PEAK = 78.6e12     # MI355X dense FP64 matrix, FLOP/s
MFMA_F32_PEAK = 157.3e12    # MI355X dense FP32 matrix, FLOP/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=["c0", "c1", "c2", "c3", "c4"],
                    help="c3: XR iteration on the 2x18-orbital dimer (the headline line); "
                         "c0: the same on the 2x9-orbital dimer; c1: density build of one "
                         "13-orbital fragment, 2-6 electrons, 32 states; c2: all blocks of six "
                         "uneven fragments, eager vs autotuned graph replay (separate metrics); "
                         "c4: XR iteration on the 4x13-orbital chain (no CPU baseline)")
    ap.add_argument("--precision", default="both", choices=["f64", "split_f32", "both"],
                    help="two-electron contraction precision; 'both' times FP64 (the value) "
                         "and split-FP32 (reported beside it with its |dE|)")
    ap.add_argument("--no-grad", action="store_true", help="skip the adjoint (diagnostics only)")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the GEMM tiles of one iteration before the warm-up (about 2%% faster; "
                         "the tuning launches then also appear in a profiler's kernel statistics)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--dump", default=None, help="write every timed launch as CSV")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def workload(name):
    from qodeapplications_amd import synthetic as syn
    if name == "c0":
        return syn.config0(seed=0), "dimer 2x9 spatial orbitals, 3/4/5 e, 8 states/block"
    if name == "c4":
        return syn.config4(seed=4), "chain 4x13 spatial orbitals, 3/4/5 e, 32 states/block"
    return syn.config3(seed=3), "dimer 2x18 spatial orbitals, 3/4/5 e, 64 states/block"


def cpu_baseline(w, budget_s):
    """Times the CPU oracle (C restatement of density_tensors.c /
    H_contractions.c, one thread) on bounded samples of the same workload
    and scales by the counted work of one iteration."""
    from oracle import oracle as orc
    from qodeapplications_amd.density import OPS_BY_CHG_DIFF
    f = w.fragments[0]
    n_orbs = f.n_orbs
    dim = 2 * n_orbs
    # (1) density build: time the ccaa-free densities of one charge pair on a
    # slice of ket determinants, as rank-1 updates per second
    ket_chg = min(f.states, key=lambda c: -f.states[c].configs.shape[1])  # most electrons
    bra_chg = ket_chg + 1
    sk, sb = f.states[ket_chg], f.states[bra_chg]
    nq = 1
    t_used = 0.0
    units = 0
    while t_used < 0.4 * budget_s and nq <= sk.configs.shape[0]:
        cfg = sk.configs[:nq]
        zk = np.ascontiguousarray(sk.coeffs[:, :nq])
        t0 = time.perf_counter()
        orc.density("caa", n_orbs, 0, sb.coeffs, zk, cfg)
        dt = time.perf_counter() - t0
        q, _, _, _ = orc.excitations("caa", n_orbs, 0, cfg)
        units = len(q) * sb.coeffs.shape[0] * zk.shape[0]
        t_used += dt
        rate_dens = units / dt
        nq *= 2
    # total rank-1 state-pair updates of one iteration (both fragments, every
    # density the blocks need, ccaa for the monomer term included as the
    # reference builds it before contracting, build_density_tensors.py:122-133)
    total_units = 0
    for fr in w.fragments:
        for bc, bs in fr.states.items():
            for kc, ks in fr.states.items():
                for op in OPS_BY_CHG_DIFF.get(bc - kc, ()):
                    n_ket = ks.configs.shape[1]
                    per_q = _paths_per_config(op, n_ket, dim)
                    total_units += per_q * ks.configs.shape[0] * bs.coeffs.shape[0] * ks.coeffs.shape[0]
    t_dens = total_units / rate_dens
    # (2) matrix elements: time dimer_ExEx calls (the dominant family)
    rng = np.random.default_rng(0)
    R1 = rng.standard_normal((dim, dim))
    R2 = rng.standard_normal((dim, dim))
    V = np.ascontiguousarray(w.integrals.V[0, 1, 0, 1])
    L = orc.lib()
    import ctypes
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    n_calls = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.4 * budget_s:
        L.orc_dimer_ExEx(ctypes.c_int64(dim), ctypes.c_int64(dim), P(R1), P(R2), P(V))
        n_calls += 1
    t_elem = (time.perf_counter() - t0) / n_calls
    spf = sum(s.coeffs.shape[0] for s in f.states.values())
    n_elem = (spf * spf) ** 2
    # every charge-conserving element costs at least one O(dim^4) call
    frac = _conserving_fraction(w)
    t_H = n_elem * frac * t_elem
    total = t_dens + t_H
    sample = (f"oracle caa density on {min(nq // 2, sk.configs.shape[0])} ket determinants "
              f"({rate_dens:.3g} rank-1 updates/s) scaled to {total_units:.3g} updates; "
              f"{n_calls} dimer_ExEx element calls ({t_elem * 1e3:.2f} ms each) scaled to "
              f"{n_elem * frac:.3g} charge-conserving pair elements; gradient not included")
    return total, sample


def _paths_per_config(op, n_ket, dim):
    """Surviving (operator-index tuple) paths per ket determinant."""
    occ, paths = n_ket, 1
    for o in reversed(op):
        if o == "a":
            paths *= occ
            occ -= 1
        else:
            paths *= dim - occ
            occ += 1
    return paths


def _conserving_fraction(w):
    f1, f2 = w.fragments[0], w.fragments[1]
    n1 = {c: s.coeffs.shape[0] for c, s in f1.states.items()}
    n2 = {c: s.coeffs.shape[0] for c, s in f2.states.items()}
    tot = sum(n1.values()) * sum(n2.values())
    good = 0
    for a in n1:
        for b in n2:
            for c in n1:
                for d in n2:
                    if a + b == c + d and abs(a - c) <= 2:
                        good += n1[a] * n2[b] * n1[c] * n2[d]
    return good / float(tot * tot)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over GPUs is "
                                   "designed but not built"}))
        sys.exit(2)
    import torch
    if args.copy_bytes:
        n = args.copy_bytes // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8}))
        return
    if args.config == "c1":
        return bench_c1(args)
    if args.config == "c2":
        return bench_c2(args)
    from qodeapplications_amd import _lib
    from qodeapplications_amd.xr import XRSystem
    from qodeapplications_amd.eigen import lowest_eigenpairs
    from qodeapplications_amd.orbitals import OrbitalRotation, energy_and_gradients

    w, wdesc = workload(args.config)
    I = w.integrals
    rot = OrbitalRotation(I.frag_orbs, I.h_sp, I.V_sp, I.U_sp)
    ctx = _lib.context(0)

    def measure(prec):
        """Warm-up, then the timed steps of one precision path; returns the
        per-step time, energy, gradient and the profiler's kernel stats."""
        X = XRSystem(w.fragments, w.integrals, w.nuc, precision=prec)

        def step():
            if args.no_grad:
                X.build_densities()
                H = X.blocks()
                M = X.xr_matrix(H=H)
                E, v = lowest_eigenpairs(M, k=1, H1=H[0])
                return float(E[0]), None
            E, gz, gk = energy_and_gradients(X, rot, lagrangian=True)
            g = torch.cat([x.reshape(-1) for x in gz.values()] + [x.reshape(-1) for x in gk])
            return E, g

        print(f"bench: workload {wdesc}; precision {prec}; warmup {args.warmup}", file=sys.stderr, flush=True)
        if args.tune:
            # tile / K-split choice per GEMM shape of one iteration (outside the timed steps)
            from qodeapplications_amd.graph import autotune
            tuned = autotune(step)
            print(f"bench: autotuned {len(tuned)} GEMM shapes", file=sys.stderr, flush=True)
        for i in range(args.warmup):
            E, g = step()
            torch.cuda.synchronize()
            print(f"bench: warmup {i} E={E:.12f}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        for i in range(args.steps):
            E, g = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.profile(False)
        if args.dump:
            ctx.profile_dump(args.dump.replace(".csv", f"_{prec}.csv"))
        ms = (t1 - t0) * 1e3 / args.steps
        kind = _lib.K_GEMM_F64 if prec == "f64" else _lib.K_GEMM_SPLIT
        stats = {}
        for name, k in (("gemm_f64", _lib.K_GEMM_F64), ("gemm_split_f32", _lib.K_GEMM_SPLIT),
                        ("gemm_grouped", _lib.K_GEMM_GROUPED), ("gemm_bandwidth", _lib.K_GEMM_BW),
                        ("other", _lib.K_OTHER),
                        ("gather", _lib.K_GATHER), ("expand", _lib.K_EXPAND),
                        ("excitation_table", _lib.K_TABLE), ("block_scatter", _lib.K_SCATTER),
                        ("braket", _lib.K_BRAKET)):
            ln, kms, fl, by = ctx.profile_stats(k)
            stats[name] = {"launches": ln, "ms_per_step": round(kms / args.steps, 3),
                           "tflops": round(fl / (kms * 1e-3) / 1e12, 2) if kms > 0 and fl > 0 else None}
        # the dominant FP64 GEMM kernel: plain tile GEMM or grouped tile GEMM
        gemm = max((ctx.profile_stats(k) + (name,) for k, name in
                    ((_lib.K_GEMM_F64, "k_dgemm"), (_lib.K_GEMM_GROUPED, "k_dgemm_grouped"))),
                   key=lambda r: r[1])
        print(f"bench: {prec} E={E:.12f} ms/step={ms:.1f} kernels={stats}", file=sys.stderr, flush=True)
        del X
        torch.cuda.empty_cache()
        return {"ms": ms, "E": E, "g": g, "stats": stats, "gemm": gemm}

    precs = ["f64", "split_f32"] if args.precision == "both" else [args.precision]
    res = {p: measure(p) for p in precs}
    main_p = precs[0]
    r0 = res[main_p]
    ms_per_step, E = r0["ms"], r0["E"]
    n, gms, gfl, gby, gname = r0["gemm"]
    avg_ms = gms / max(n, 1)
    achieved = gfl / (gms * 1e-3) if gms > 0 else 0.0
    stats = r0["stats"]
    gn = None if r0["g"] is None else float(torch.linalg.norm(r0["g"]))
    split = None
    if "split_f32" in res and main_p == "f64":
        rs = res["split_f32"]
        split = {"s_per_iter": rs["ms"] / 1e3, "E0": rs["E"], "abs_dE_vs_f64": abs(rs["E"] - E),
                 "speedup_vs_f64": ms_per_step / rs["ms"], "kernels": rs["stats"]}
        if rs["g"] is not None and r0["g"] is not None:
            split["grad_max_abs_err_vs_f64"] = float((rs["g"] - r0["g"]).abs().max())
            split["grad_rel_err_vs_f64"] = float(torch.linalg.norm(rs["g"] - r0["g"]) / torch.linalg.norm(r0["g"]))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "c3_dimer_traffic.json")
    if args.config == "c3" and os.path.exists(tfile):
        # HBM bytes per FP64 GEMM launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
        # on this workload (the profile summary names the commands)
        traffic = json.load(open(tfile))["traffic_bytes_per_launch"].get(gname)

    cpu = None
    if not args.no_cpu and len(w.fragments) == 2:
        t_cpu, sample = cpu_baseline(w, args.cpu_seconds)
        cpu = {"value": t_cpu, "unit": "s/iter", "cores": 1, "kind": "port", "sample": sample}

    out = {
        "metric": "XR H-build+grad s/iter & % roofline: FP64 plain vs split-FP32, |dE|<=1e-9 Eh",
        "value": ms_per_step / 1e3,
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random orthonormal CI vectors, 8-fold symmetric V)",
        "config": {"workload": wdesc, "precision_path": main_p, "gradient": not args.no_grad,
                   "E0": E, "grad_norm": gn},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": MFMA_F64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / MFMA_F64_PEAK,
                     "kernel": f"{gname} (FP64 MFMA GEMM)", "launches_per_step": n / args.steps,
                     "avg_launch_ms": avg_ms, "traffic": traffic,
                     "algorithmic_bytes_per_launch": gby / max(n, 1)},
        "kernels": stats,
        "split_fp32": split,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_c1(args):
    """Single-fragment density build at scale: every 1-4 index transition
    density of a 13-spatial-orbital fragment with 2-6 electrons and 32 states
    per charge block (42.4 GB of FP64 tensors).  One step = the whole build."""
    import torch
    from qodeapplications_amd import _lib, synthetic as syn
    from qodeapplications_amd import density as D
    frag = syn.config1_fragment(seed=1)
    blocks = D.charge_blocks(frag.states, frag.n_orbs, frag.n_core, device=0)
    dim = 2 * frag.n_orbs
    flops = 0.0          # 2 * nnz * nI * nJ, the reference loop count
    out_bytes = 0
    for bc, bb in blocks.items():
        for kc, kb in blocks.items():
            for op in D.OPS_BY_CHG_DIFF_FULL.get(bc - kc, ()):
                flops += 2.0 * _paths_per_config(op, kb.n_elec, dim) * kb.n_configs * bb.n_states * kb.n_states
                out_bytes += 8 * bb.n_states * kb.n_states * dim ** len(op)
    in_bytes = sum(8 * b.coeffs.numel() for b in blocks.values())
    ctx = _lib.context(0)
    wdesc = "fragment 13 spatial orbitals, 2-6 e, 32 states/block, all 1-4 index densities"
    print(f"bench: workload {wdesc}; warmup {args.warmup}", file=sys.stderr, flush=True)

    def step():
        rho = D.build_density_arrays(blocks, D.OPS_BY_CHG_DIFF_FULL, skip=())
        del rho

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"bench: warmup {i} done", file=sys.stderr, flush=True)
    ctx.profile(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.profile(False)
    if args.dump:
        ctx.profile_dump(args.dump)
    s = (t1 - t0) / args.steps
    n, gms, gfl, gby = ctx.profile_stats(_lib.K_GEMM_F64)
    stats = {}
    for name, k in (("gemm_f64", _lib.K_GEMM_F64), ("gather", _lib.K_GATHER), ("expand", _lib.K_EXPAND),
                    ("excitation_table", _lib.K_TABLE)):
        ln, kms, fl, by = ctx.profile_stats(k)
        stats[name] = {"launches": ln, "ms_per_step": round(kms / args.steps, 3),
                       "tflops": round(fl / (kms * 1e-3) / 1e12, 2) if kms > 0 and fl > 0 else None}
    out = {
        "metric": "single-fragment density build s/build",
        "value": s, "unit": "s/build", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": s * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded random orthonormal CI vectors)",
        "config": {"workload": wdesc, "reference_loop_flops": flops, "output_bytes": out_bytes},
        "reference_loop_tflops": flops / s / 1e12,
        "output_gbs": (out_bytes + in_bytes) / s / 1e9,
        "roofline": {"bound": "mfma", "achieved": gfl / (gms * 1e-3) / 1e12 if gms else 0.0,
                     "peak": MFMA_F64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": gfl / (gms * 1e-3) / MFMA_F64_PEAK if gms else 0.0,
                     "kernel": "k_dgemm (FP64 MFMA GEMM)", "launches_per_step": n / args.steps,
                     "avg_launch_ms": gms / max(n, 1), "traffic": None},
        "kernels": stats,
    }
    print(json.dumps(out))


def bench_c2(args):
    """All densities, monomer and pair blocks (15 pairs) of six fragments of
    9, 9, 13, 13, 18, 18 spatial orbitals with 3/4/5 electrons and 8-64
    states per charge block.  One step = one full block build; timed eagerly
    and as replays of the autotuned, graph-captured build."""
    import torch
    from qodeapplications_amd import _lib, synthetic as syn
    from qodeapplications_amd.graph import CapturedBuild
    from qodeapplications_amd.xr import XRSystem
    w = syn.config2(seed=2)
    X = XRSystem(w.fragments, w.integrals, w.nuc)
    ctx = _lib.context(0)
    wdesc = ("six fragments 9,9,13,13,18,18 spatial orbitals, 3/4/5 e, states/block "
             + ",".join(str(next(iter(f.states.values())).coeffs.shape[0]) for f in w.fragments)
             + ", all 15 pair blocks")
    print(f"bench: workload {wdesc}", file=sys.stderr, flush=True)

    def eager():
        X.build_densities()
        return X.blocks()

    for i in range(args.warmup):
        eager()
        torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        H = eager()
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / args.steps
    ctx.profile(False)
    n, gms, gfl, gby = ctx.profile_stats(_lib.K_GEMM_F64)
    del H
    print(f"bench: eager {t_eager * 1e3:.1f} ms/build", file=sys.stderr, flush=True)
    cap = CapturedBuild(X, tune=True)
    for i in range(args.warmup):
        cap.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        cap.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / args.steps
    print(f"bench: graph {t_graph * 1e3:.1f} ms/build, {len(cap.tuned)} tuned shapes",
          file=sys.stderr, flush=True)
    out = {
        "metric": "XR block build s/build (6 uneven fragments, 15 pair blocks)",
        "value": t_graph, "unit": "s/build", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_graph * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded random orthonormal CI vectors, 8-fold symmetric V)",
        "config": {"workload": wdesc},
        "eager_s": t_eager, "graph_speedup": t_eager / t_graph,
        "tuned_shapes": [{"shape": list(k), "tile": v[0], "splits": v[1], "tuned_ms": v[2], "untuned_ms": v[3]}
                         for k, v in sorted(cap.tuned.items(), key=lambda kv: -kv[1][3])[:12]],
        "roofline": {"bound": "mfma", "achieved": gfl / (gms * 1e-3) / 1e12 if gms else 0.0,
                     "peak": MFMA_F64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": gfl / (gms * 1e-3) / MFMA_F64_PEAK if gms else 0.0,
                     "kernel": "k_dgemm (FP64 MFMA GEMM), eager build", "launches_per_step": n / args.steps,
                     "avg_launch_ms": gms / max(n, 1), "traffic": None},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
