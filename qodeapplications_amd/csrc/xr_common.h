// Shared definitions for the XR HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>
#include <array>
#include <tuple>
#include <mutex>

#include "../../include/xr_hip.h"

#define XR_MAX_SO 64      // spin orbitals per fragment: one 64-bit occupation word
#define XR_MAX_OPS 5

// ---------------------------------------------------------------------------
// binomial table in constant memory: C(n, k) for 0 <= n, k <= 64
// ---------------------------------------------------------------------------
extern __constant__ uint64_t c_binom[65 * 65];

__device__ __forceinline__ uint64_t d_binom(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    return c_binom[n * 65 + k];
}

__device__ __forceinline__ int d_popc(uint64_t x) { return __popcll(x); }

// Parity of the occupied spin orbitals strictly above p: the sign that
// inserting or removing p picks up in an ascending occupation list
// (general-XRCC/density_tensors.c:80-113).
__device__ __forceinline__ int d_sign_above(uint64_t occ, int p) {
    uint64_t above = (p >= 63) ? 0ull : (occ >> (p + 1));
    return (d_popc(above) & 1) ? -1 : 1;
}

// Lexical rank of a k-subset of {0..n-1} given as a bit mask:
//   rank = C(n,k) - 1 - sum_i C(n-1-c_i, k-i),  c_0 < c_1 < ... the set bits.
__device__ __forceinline__ int64_t d_lex_rank(uint64_t v, int n) {
    int k = d_popc(v);
    int64_t acc = 0;
    int i = 0;
    while (v) {
        int c = __ffsll((unsigned long long)v) - 1;
        v &= v - 1;
        acc += (int64_t)d_binom(n - 1 - c, k - i);
        i++;
    }
    return (int64_t)d_binom(n, k) - 1 - acc;
}

// Inverse of d_lex_rank: the k-subset of {0..n-1} with the given rank.
__device__ __forceinline__ uint64_t d_lex_unrank(int64_t rank, int n, int k) {
    int64_t s = (int64_t)d_binom(n, k) - 1 - rank;
    uint64_t v = 0;
    int d = n - 1;
    for (int i = 0; i < k; i++) {
        int m = k - i;
        while (d >= 0 && (int64_t)d_binom(d, m) > s) d--;
        // d_i = d  ->  c_i = n-1-d
        s -= (int64_t)d_binom(d, m);
        v |= 1ull << (n - 1 - d);
        d--;
    }
    return v;
}

// Valence compression: drop the n_core lowest alpha and beta spin orbitals.
__device__ __forceinline__ uint64_t d_to_valence(uint64_t occ, int n_orbs, int n_core) {
    if (n_core == 0) return occ;
    int nv = n_orbs - n_core;
    uint64_t m = (nv >= 64) ? ~0ull : ((1ull << nv) - 1);
    uint64_t a = (occ >> n_core) & m;
    uint64_t b = (occ >> (n_orbs + n_core)) & m;
    return a | (b << nv);
}

__device__ __forceinline__ uint64_t d_core_mask(int n_orbs, int n_core) {
    uint64_t m = (n_core >= 64) ? ~0ull : ((1ull << n_core) - 1);
    return m | (m << n_orbs);
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// kernel classes timed by the profiler (xr_profile / xr_profile_stats)
// XR_K_GEMM_F64: the MFMA tile GEMM kernel alone; XR_K_GEMM_BW: the
// bandwidth-bound tall-skinny product kernels; XR_K_GEMM_GROUPED: the
// grouped tile GEMM (many products in one grid); XR_K_GEMM_WORKLIST: the
// pair program's work-list GEMM; XR_K_GEMM_SPLIT_I8: the split-int8 GEMM
// kernel; XR_K_I8_SLICE: its exponent and digit-plane passes; reductions of
// staged GEMMs count as XR_K_OTHER.
enum { XR_K_GEMM_F64 = 0, XR_K_GATHER = 1, XR_K_EXPAND = 2, XR_K_TABLE = 3, XR_K_SCATTER = 4,
       XR_K_BRAKET = 5, XR_K_OTHER = 6, XR_K_GEMM_SPLIT = 7, XR_K_GEMM_BW = 8,
       XR_K_GEMM_GROUPED = 9, XR_K_GEMM_WORKLIST = 10, XR_K_GEMM_SPLIT_I8 = 11, XR_K_I8_SLICE = 12,
       XR_K_COUNT = 13 };

struct ProfRec {
    int kind;
    int64_t m = 0, n = 0, k = 0, batch = 0;
    int ta = -1, tb = -1;
    double flops, bytes;
    hipEvent_t a, b;
};

struct xr_ctx {
    int device = 0;
    int precision = XR_PREC_F64;  // two-electron contraction GEMMs
    bool prof = false;
    uint32_t prof_kinds = 0xffffffffu;  // kernel classes the profiler records (xr_profile_kinds)
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> ev_pool;
    hipStream_t stream = nullptr;
    std::string err;
    // growable scratch buffers (device)
    std::vector<void*> ws;
    std::vector<size_t> ws_size;
    size_t ws_limit = 0;        // cap on the workspaces' total bytes (0: none; xr_set_memory_limit)
    // cache of K-space occupation masks keyed by (dim, n_elec)
    std::map<std::pair<int, int>, std::pair<uint64_t*, int64_t>> kspace;
    // side-adjoint index tables (one int32 code per source determinant and
    // tuple), keyed by the side string and the spaces it maps between
    std::map<std::array<int64_t, 12>, std::pair<int32_t*, int64_t>> sidetab;
    // GEMM autotune table: (ta, tb, M, N, K, batch) -> (tile, splits)
    std::map<std::array<int64_t, 6>, std::pair<int, int>> tune;
    // captured graphs holding the workspace addresses (xr_pin): while > 0 a
    // workspace slot may not be reallocated
    int pins = 0;
    bool pin_blocked = false;   // the last xr_workspace refused to grow a pinned slot
    // the caller guarantees V[pqrs] = V[rspq] for the monomer sigma adjoint
    // (xr_assume_hermitian_v): enables its one-gather path for a block with itself
    bool hermitian_v = false;
    // no-allocation calls (xr_gemm_f64_noalloc): the device address of the
    // library's static scratch array on this context's device, its size, and
    // whether a call is running in that mode (xr_workspace then hands out
    // the scratch array instead of allocating)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    bool noalloc = false;
    bool noalloc_blocked = false;
    // pinned staging of xr_pair_blocks' descriptor tables and the event that
    // marks the last copy out of it
    void* pair_host = nullptr;
    size_t pair_host_bytes = 0;
    hipEvent_t pair_copied = nullptr;
    std::mutex mu;
};

int xr_fail(xr_ctx* c, int code, const std::string& msg);

// Brackets one kernel launch with HIP events on the context stream when the
// profiler is on (kind: XR_K_*; flops / bytes: algorithmic work of the launch).
struct ProfScope {
    xr_ctx* c;
    int idx = -1;
    ProfScope(xr_ctx* c_, int kind, double flops, double bytes);
    void shape(int64_t m, int64_t n, int64_t k, int64_t b, int ta = -1, int tb = -1);
    ~ProfScope();
};
int xr_hip_check(xr_ctx* c, hipError_t e, const char* what);
void* xr_workspace(xr_ctx* c, int slot, size_t bytes);
int xr_noalloc_init(xr_ctx* c);
int xr_fill(xr_ctx* c, double* p, int64_t n, double v);

#define XR_HIP(c, expr)                                         \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return xr_hip_check((c), _e, #expr); \
    } while (0)

#define XR_LAUNCH_CHECK(c) XR_HIP(c, hipGetLastError())

static inline int64_t h_binom(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    int64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Elementwise kernels loop with a grid stride over a capped grid: a launch
// may not exceed 2^32 work-items per grid dimension on this hardware.
#define XR_GRID_STRIDE(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (int64_t)gridDim.x * blockDim.x)

static inline unsigned xr_grid(int64_t n, int bs = 256) {
    int64_t b = ceil_div(n, bs);
    if (b < 1) b = 1;
    return (unsigned)(b < (1 << 20) ? b : (1 << 20));
}
