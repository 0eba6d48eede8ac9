"""ctypes binding of libxrhip.so (the C ABI in include/xr_hip.h).

The HIP library is the only compute path: if it is missing or no GPU is
visible, every entry point raises instead of falling back to a CPU path.
"""
import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxrhip.so")

PREC_F64, PREC_SPLIT_F32, PREC_SPLIT_I8 = 0, 1, 2
(K_GEMM_F64, K_GATHER, K_EXPAND, K_TABLE, K_SCATTER, K_BRAKET, K_OTHER, K_GEMM_SPLIT, K_GEMM_BW,
 K_GEMM_GROUPED, K_GEMM_WORKLIST, K_GEMM_SPLIT_I8, K_I8_SLICE) = range(13)
XR_OK, XR_EINVAL, XR_ESHAPE, XR_EINVAL_CONFIG_ORDER, XR_EPAULI, XR_ENOMEM, XR_EHIP, XR_EUNSUPPORTED, \
    XR_EBUSY = range(9)
_CODE_NAMES = {1: "XR_EINVAL", 2: "XR_ESHAPE", 3: "XR_EINVAL_CONFIG_ORDER", 4: "XR_EPAULI",
               5: "XR_ENOMEM", 6: "XR_EHIP", 7: "XR_EUNSUPPORTED", 8: "XR_EBUSY"}

# every symbol include/xr_hip.h declares, with its ctypes signature
_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double
_S = ctypes.c_char_p
SIGNATURES = {
    "xr_create": (_I, [_I, _P, ctypes.POINTER(_P)]),
    "xr_destroy": (_I, [_P]),
    "xr_set_stream": (_I, [_P, _P]),
    "xr_last_error": (_S, [_P]),
    "xr_synchronize": (_I, [_P]),
    "xr_pin": (_I, [_P, _I]),
    "xr_trim": (_I, [_P]),
    "xr_set_memory_limit": (_I, [_P, _L]),
    "xr_assume_hermitian_v": (_I, [_P, _I]),
    "xr_profile": (_I, [_P, _I]),
    "xr_profile_stats": (_I, [_P, _I, ctypes.POINTER(_L), ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "xr_profile_dump": (_I, [_P, _S]),
    "xr_configs_to_masks": (_I, [_P, _P, _L, _I, _I, _I, _P]),
    "xr_config_rank": (_I, [_P, _P, _L, _I, _I, _P]),
    "xr_excitation_count": (_I, [_P, _S, _I, _I, _I, _P, _L, _L, _P, ctypes.POINTER(_L)]),
    "xr_excitation_fill": (_I, [_P, _S, _I, _I, _I, _P, _L, _L, _P, _P, _P, _P, _P]),
    "xr_density": (_I, [_P, _S, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P]),
    "xr_monomer_2e": (_I, [_P, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P, _P]),
    "xr_density_adjoint": (_I, [_P, _S, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P, _P, _P]),
    "xr_monomer_2e_adjoint": (_I, [_P, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P, _P, _P, _P, _P]),
    "xr_monomer_2e_keep": (_I, [_P, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P, _P, _P, _L]),
    "xr_monomer_2e_adjoint_keep": (_I, [_P, _I, _I, _I, _L, _L, _P, _L, _L, _P, _P, _P, _P, _P, _P, _P, _L]),
    "xr_pair_outer": (_I, [_P, _P, _L, _P, _P, _P, _P, _L, _L, _L, _L, _I, _P]),
    "xr_gemm_tune": (_I, [_P, _I, _I, _L, _L, _L, _L, _I, _P, _P, _P]),
    "xr_gemm_tune_reset": (_I, [_P]),
    "xr_symmetrize": (_I, [_P, _P, _L]),
    "xr_conjugate": (_I, [_P, _P, _L, _L, _I, _I, _P, _I]),
    "xr_gemm_grouped": (_I, [_P, _I, _I, _I, _P, _P, _P, _D, _P, _P, _P, _P, _P, _P]),
    "xr_axes_pack": (_I, [_P, _I, _P, _L, _I, _I, _P, _L, _P, _P, _L, _I, _D]),
    "xr_axis_apply_multi": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _D]),
    "xr_axes_pack_multi": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _D]),
    "xr_pair_gather": (_I, [_P, _P, _L, _L, _L, _P, _P, _P, _P, _P, _L, _L, _L, _L, _I, _P]),
    "xr_gemm_f64": (_I, [_P, _I, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _L]),
    "xr_gemm_split_f32": (_I, [_P, _I, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _L]),
    "xr_gemm_split_i8": (_I, [_P, _I, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _L]),
    "xr_gemm_syrk": (_I, [_P, _I, _L, _L, _D, _P, _L, _P, _L]),
    "xr_gemm_mode": (_I, [_P, _I, _I, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _L]),
    "xr_precond_divide": (_I, [_P, _L, _L, _P, _L, _P, _L, _P, _P, _D, _D]),
    "xr_residual": (_I, [_P, _L, _I, _I, _P, _L, _P, _L, _P, _P, _L, _P]),
    "xr_column_norms": (_I, [_P, _L, _I, _P, _L, _P, _I]),
    "xr_transpose": (_I, [_P, _L, _L, _P, _L, _P, _L]),
    "xr_sector_diag": (_I, [_P, _I, _I, _P, _P, _P, _P, _P]),
    "xr_column_dots": (_I, [_P, _L, _L, _P, _L, _P, _L, _P]),
    "xr_spin_blocks": (_I, [_P, _I, _P, _P, _P, _P, _P, _L]),
    "xr_spin_blocks_adjoint": (_I, [_P, _I, _P, _P, _P, _P, _P, _L]),
    "xr_permute4": (_I, [_P, _P, _P, _P, _P, _I]),
    "xr_pair_dress": (_I, [_P, _L, _P, _L, _P, _P, _P, _L, _P, _L, _D, _D, _P, _L]),
    "xr_lanczos": (_I, [_P, _I, _P, _P, _P, _I, _P, _P, _P]),
    "xr_add_diagonal": (_I, [_P, _I, _P, _P, _P, _P]),
    "xr_cholesky_test": (_I, [_P, _I, _P, _P, _P, _P]),
    "xr_gemm_f64_noalloc": (_I, [_P, _I, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _L]),
    "xr_set_precision": (_I, [_P, _I]),
    "xr_profile_kinds": (_I, [_P, ctypes.c_uint32]),
    "xr_monomer_1e": (_I, [_P, _L, _L, _P, _P, _P]),
    "xr_dimer_2min2pls": (_I, [_P, _L, _L, _L, _P, _L, _P, _P, _P]),
    "xr_dimer_1min1pls_1e": (_I, [_P, _L, _L, _L, _P, _L, _P, _P, _P]),
    "xr_dimer_1min1pls_2e": (_I, [_P, _L, _L, _L, _P, _P, _L, _P, _P, _P, _P, _P]),
    "xr_dimer_1min1pls": (_I, [_P, _L, _L, _L, _P, _P, _L, _P, _P, _P, _P, _P, _P]),
    "xr_trimer_2min1pls1pls": (_I, [_P, _L, _L, _L, _L, _P, _L, _P, _L, _P, _P, _P]),
    "xr_trimer_2pls1min1min": (_I, [_P, _L, _L, _L, _L, _P, _L, _P, _L, _P, _P, _P]),
    "xr_trimer_Ex1min1pls": (_I, [_P, _L, _L, _L, _L, _P, _L, _P, _L, _P, _P, _P]),
    "xr_block_scatter": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _L]),
    "xr_braket_assemble": (_I, [_P, _I, _L, _P, _P, _P, _P, _P]),
    "xr_dimer_ExEx": (_I, [_P, _L, _L, _L, _P, _L, _P, _P, _P]),
    "xr_pair_blocks": (_I, [_P, _P, _I, _P, _P, _P, _I, _I]),
    "xr_pair_plan_create": (_I, [_P, _P, _I, _P, _P, _P, _I, _I, ctypes.POINTER(_P)]),
    "xr_pair_plan_run": (_I, [_P]),
    "xr_pair_plan_info": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_L), ctypes.POINTER(_D),
                               ctypes.POINTER(_L)]),
    "xr_pair_plan_destroy": (_I, [_P]),
    "xr_graph_begin": (_I, [_P]),
    "xr_graph_end": (_I, [_P, ctypes.POINTER(_P)]),
    "xr_graph_launch": (_I, [_P]),
    "xr_graph_destroy": (_I, [_P]),
}

PAIR_2MIN2PLS, PAIR_1MIN1PLS, PAIR_EXEX = 0, 1, 2


class PairJob(ctypes.Structure):
    """struct xr_pair_job of include/xr_hip.h (one charge combination of one
    fragment pair for xr_pair_blocks / xr_pair_plan_create)."""
    _fields_ = [("kind", ctypes.c_int32), ("transposed", ctypes.c_int32),
                ("n1", _L), ("n2", _L), ("nI1", _L), ("nJ1", _L), ("nI2", _L), ("nJ2", _L),
                ("R1", _P * 2), ("R2", _P * 2), ("V", _P * 2), ("h", _P * 2),
                ("scale", _D), ("nuc", _D), ("out", _P), ("ldo", _L), ("mul", _L),
                ("rows1", _P), ("cols1", _P), ("rows2", _P), ("cols2", _P)]


class XRError(RuntimeError):
    """Raised when a libxrhip entry point returns a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"{_CODE_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def build(force=False):
    """Compile libxrhip.so for gfx950 with the package Makefile."""
    csrc = os.path.join(_HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", csrc], check=True)


def load():
    """The loaded library (raises if it was not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                                  "(the HIP library is required; there is no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


class Context:
    """Owns one xr_ctx on a device, bound to torch's current stream there."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: libxrhip needs an MI355X (there is no CPU fallback)")
        self.L = load()
        self.device = device
        with torch.cuda.device(device):
            stream = torch.cuda.current_stream(device).cuda_stream
        h = _P()
        rc = self.L.xr_create(device, _P(stream), ctypes.byref(h))
        if rc != XR_OK:
            raise XRError(rc, "xr_create failed")
        self.h = h
        self.precision = PREC_F64

    def check(self, rc):
        if rc != XR_OK:
            raise XRError(rc, self.L.xr_last_error(self.h).decode())

    def call(self, name, *args):
        self.check(getattr(self.L, name)(self.h, *args))

    def sync_stream(self):
        import torch
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.check(self.L.xr_set_stream(self.h, _P(stream)))

    def set_precision(self, mode):
        """PREC_F64, PREC_SPLIT_F32 or PREC_SPLIT_I8 for the two-electron contraction GEMMs."""
        self.check(self.L.xr_set_precision(self.h, int(mode)))
        self.precision = int(mode)

    def profile(self, enable):
        self.check(self.L.xr_profile(self.h, 1 if enable else 0))

    def profile_kinds(self, kinds=None):
        """Records only the listed kernel classes (None: all)."""
        mask = 0xffffffff if kinds is None else sum(1 << int(k) for k in kinds)
        self.check(self.L.xr_profile_kinds(self.h, ctypes.c_uint32(mask)))

    def profile_stats(self, kind):
        """(launches, total_ms, flops, bytes) of one kernel class since profile(True)."""
        n, ms, fl, by = _L(0), _D(0), _D(0), _D(0)
        self.check(self.L.xr_profile_stats(self.h, kind, ctypes.byref(n), ctypes.byref(ms),
                                           ctypes.byref(fl), ctypes.byref(by)))
        return n.value, ms.value, fl.value, by.value

    def profile_dump(self, path):
        self.check(self.L.xr_profile_dump(self.h, path.encode()))

    def close(self):
        if getattr(self, "h", None):
            self.L.xr_destroy(self.h)
            self.h = None


_contexts = {}


def context(device=0):
    """Per-device singleton context, re-bound to torch's current stream."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = _contexts[device] = Context(device)
    ctx.sync_stream()
    return ctx


def ptr(t):
    """Device pointer of a torch tensor (None for None)."""
    return None if t is None else _P(t.data_ptr())
