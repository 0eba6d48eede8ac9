"""Batched pair blocks and graph capture on libxrhip (xr_pair_blocks,
xr_pair_plan_*, xr_graph_*).

A list of PairJob records (one per charge combination of a fragment pair,
built by XRSystem) becomes one program of a few launches: the permuted
integral blocks, every density x integral product as a work-list GEMM, the
final products, one scatter into the output matrices and their A + A^T.
pair_blocks runs it once on the context's workspaces; PairPlan owns its
intermediates, so run() allocates nothing and can be captured (Graph, or
torch's graph capture) and replayed.
"""
import ctypes

from ._lib import PairJob, context, XR_OK, XRError

_P = ctypes.c_void_p
_L = ctypes.c_int64


def _outs(outs):
    n = len(outs)
    for o in outs:
        if o.dim() != 2 or o.shape[0] != o.shape[1] or o.stride(1) != 1:
            raise ValueError("pair outputs must be square row-major matrices")
    ptrs = (_P * max(n, 1))(*[o.data_ptr() for o in outs])
    sizes = (_L * max(n, 1))(*[o.shape[0] for o in outs])
    lds = (_L * max(n, 1))(*[o.stride(0) for o in outs])
    return ptrs, sizes, lds, n


def _jobs(jobs):
    arr = (PairJob * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = j
    return arr, len(jobs)


def pair_blocks(jobs, outs, mode=0, device=0):
    """Zeroes outs, adds every job's block and symmetrises each output
    (xr_pair_blocks, on torch's current stream)."""
    ctx = context(device)
    arr, n = _jobs(jobs)
    p, s, l, no = _outs(outs)
    ctx.check(ctx.L.xr_pair_blocks(ctx.h, arr, n, p, s, l, no, int(mode)))


class PairPlan:
    """xr_pair_plan over jobs and outs; keep: the tensors the jobs point
    into (held for the plan's lifetime)."""

    def __init__(self, jobs, outs, mode=0, device=0, keep=()):
        self.ctx = context(device)
        self.outs = list(outs)
        self.keep = list(keep)
        arr, n = _jobs(jobs)
        p, s, l, no = _outs(self.outs)
        h = _P()
        self.ctx.check(self.ctx.L.xr_pair_plan_create(self.ctx.h, arr, n, p, s, l, no, int(mode), ctypes.byref(h)))
        self.h = h

    def run(self):
        """One run on the context's current stream (torch's current stream
        when called through context())."""
        if self.h is None:
            raise RuntimeError("the pair plan was closed")
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.xr_pair_plan_run(self.h))
        return self.outs

    def info(self):
        """{launches, workgroups, flops, arena_bytes} of one run."""
        a, b, c, d = _L(0), _L(0), ctypes.c_double(0), _L(0)
        self.ctx.check(self.ctx.L.xr_pair_plan_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                    ctypes.byref(d)))
        return {"launches": a.value, "workgroups": b.value, "flops": c.value, "arena_bytes": d.value}

    def close(self):
        if getattr(self, "h", None) is not None:
            self.ctx.L.xr_pair_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """HIP graph of the calls made on a context between begin() and end()
    (xr_graph_*), replayed with launch().  The context must run on a stream
    of its own (e.g. inside `with torch.cuda.stream(s)` around context()),
    and the captured sequence must have run once before (its workspaces are
    pinned while the graph exists)."""

    def __init__(self, device=0):
        self.ctx = context(device)
        self.h = None

    def begin(self):
        self.ctx.check(self.ctx.L.xr_graph_begin(self.ctx.h))

    def end(self):
        h = _P()
        self.ctx.check(self.ctx.L.xr_graph_end(self.ctx.h, ctypes.byref(h)))
        self.h = h

    def launch(self):
        if self.h is None:
            raise RuntimeError("no captured graph")
        self.ctx.check(self.ctx.L.xr_graph_launch(self.h))

    def close(self):
        if self.h is not None:
            rc = self.ctx.L.xr_graph_destroy(self.h)
            self.h = None
            if rc != XR_OK:
                raise XRError(rc, "xr_graph_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
