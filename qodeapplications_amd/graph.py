"""Autotuned, graph-captured build of the XR blocks.

The reference evaluates every matrix element through a Python call into C
(general-XRCC/build_H.py:37-188, driven element by element in
general-XRCC/test_H.py:49-138).  Here one build of an XRSystem (all
densities, the monomer sigma terms, every monomer, pair and optionally
trimer block) is a fixed sequence of libxrhip launches on one stream, so it
is captured once into a HIP graph and replayed as a single launch:

* autotune(fn): runs fn once with the profiler on, collects the shapes of
  its FP64 GEMMs and times the tile / K-split candidates of each shape
  (xr_gemm_tune); later launches of those shapes use the winner;
* CapturedBuild(system): warm-up (workspaces and caches reach their final
  sizes, so the captured sequence allocates and synchronises nothing), a
  pair plan (every pair matrix as one batched program on fixed addresses:
  the densities are rebuilt in place), then capture of build_densities +
  blocks; replay() recomputes every block for the coefficients currently
  held in the system's coefficient tensors (update them in place with
  set_coeffs).
"""
import csv
import ctypes
import os
import tempfile

import torch

from ._lib import K_GEMM_F64, context
from .xr import precision


def gemm_shapes(fn, device=0):
    """{(transA, transB, M, N, K, batch): [launches, total_ms]} of the FP64
    GEMMs fn() launches."""
    ctx = context(device)
    ctx.profile(True)
    try:
        fn()
        torch.cuda.synchronize()
        fd, path = tempfile.mkstemp(suffix=".csv")
        os.close(fd)
        ctx.profile_dump(path)
    finally:
        ctx.profile(False)
    shapes = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if int(row["kind"]) != K_GEMM_F64 or int(row["ta"]) < 0:
                continue
            key = (int(row["ta"]), int(row["tb"]), int(row["m"]), int(row["n"]), int(row["k"]),
                   int(row["batch"]))
            s = shapes.setdefault(key, [0, 0.0])
            s[0] += 1
            s[1] += float(row["ms"])
    os.unlink(path)
    return shapes


def autotune(fn, device=0, reps=3, min_ms=0.05):
    """Tunes every FP64 GEMM shape of fn() that takes at least min_ms in
    total.  Returns {shape: (tile, splits, tuned_ms, untuned_ms)}."""
    ctx = context(device)
    out = {}
    for key, (n, ms) in sorted(gemm_shapes(fn, device).items(), key=lambda kv: -kv[1][1]):
        if ms < min_ms:
            continue
        ta, tb, M, N, K, batch = key
        if 8 * (M * K + K * N + M * N) * batch > (16 << 30):
            continue
        tile, splits, best = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        r = max(1, min(reps, int(5.0 / max(ms / n, 1e-3))))     # fewer repeats for long launches
        ctx.call("xr_gemm_tune", ta, tb, M, N, K, batch, r, ctypes.byref(tile), ctypes.byref(splits),
                 ctypes.byref(best))
        if tile.value >= 0:
            out[key] = (tile.value, splits.value, best.value, ms / n)
    return out


class CapturedBuild:
    """Densities and blocks of an XRSystem as one replayable HIP graph."""

    def __init__(self, system, with_trimers=False, tune=True):
        self.X = system
        dev = system.device
        self.ctx = context(dev.index)

        def eager():
            system.build_densities(reuse=True)
            with precision(system.precision, dev.index):
                return system._blocks(with_trimers)

        self.tuned = autotune(eager, dev.index) if tune else {}
        eager()                                   # densities at their final addresses
        nf = len(system.frags)
        self.plan, _ = system.pair_plan([(m, n) for m in range(nf) for n in range(m + 1, nf)])

        def build():
            system.build_densities(reuse=True)
            with precision(system.precision, dev.index):
                return system._blocks(with_trimers, pair_plan=self.plan)

        build()                                   # final workspace sizes, warm caches
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        self._pinned = False
        try:
            with torch.cuda.graph(self.graph):
                self.ctx.sync_stream()            # launch on the capture stream
                self.H = build()
        finally:
            self.ctx.sync_stream()                # back to the ambient stream
        torch.cuda.synchronize()
        # the graph replays the workspace addresses it was captured with: no
        # call on this context may reallocate them while it exists
        self.ctx.call("xr_pin", 1)
        self._pinned = True

    def release(self):
        """Drops the graph and lets the context's workspaces grow again."""
        if getattr(self, "_pinned", False):
            self._pinned = False
            self.graph = None
            self.ctx.call("xr_pin", -1)
            self.plan.close()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def set_coeffs(self, coeffs):
        """coeffs: {(fragment, charge): (n_states x n_configs) tensor}; copied
        into the tensors the graph reads."""
        for (m, c), z in coeffs.items():
            self.X.frags[m].blocks[c].coeffs.copy_(z)

    def replay(self):
        """Recomputes every density and block; returns (H1, H2, H3) (the same
        tensors on every replay)."""
        if self.graph is None:
            raise RuntimeError("the captured build was released")
        self.graph.replay()
        return self.H
