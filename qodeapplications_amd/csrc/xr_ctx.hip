// Context, errors, workspaces, binomial table, configuration masks and ranks.
#include "xr_common.h"
#include <cstring>

__constant__ uint64_t c_binom[65 * 65];

int xr_fail(xr_ctx* c, int code, const std::string& msg) {
    if (c && c->noalloc_blocked && code == XR_ENOMEM) {
        c->noalloc_blocked = false;
        c->err = msg + ": the call needs more staging memory than the no-allocation scratch holds (" +
                 std::to_string(c->scratch_bytes >> 20) + " MiB); use the allocating entry point";
        return XR_EINVAL;
    }
    if (c && c->pin_blocked && code == XR_ENOMEM) {
        // the allocation failed because a captured graph pins the workspaces
        c->pin_blocked = false;
        c->err = msg + ": a workspace would have to grow while a captured graph holds its address "
                       "(xr_pin); release the capture or capture after the largest call";
        return XR_EBUSY;
    }
    if (c) c->err = msg;
    return code;
}

int xr_hip_check(xr_ctx* c, hipError_t e, const char* what) {
    std::string m = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return xr_fail(c, e == hipErrorOutOfMemory ? XR_ENOMEM : XR_EHIP, m);
}

// Returns a device buffer of at least `bytes` for scratch slot `slot`,
// growing it when needed (the old contents are not kept).
void* xr_workspace(xr_ctx* c, int slot, size_t bytes) {
    if (c->noalloc) {           // a no-allocation call: the static scratch array or nothing
        if (c->scratch && bytes <= c->scratch_bytes) return c->scratch;
        c->noalloc_blocked = true;
        return nullptr;
    }
    if ((int)c->ws.size() <= slot) {
        c->ws.resize(slot + 1, nullptr);
        c->ws_size.resize(slot + 1, 0);
    }
    if (c->ws_size[slot] >= bytes && c->ws[slot]) return c->ws[slot];
    if (c->ws_limit) {          // the caller's cap on the workspaces' total
        size_t total = 0;
        for (size_t i = 0; i < c->ws_size.size(); i++)
            if ((int)i != slot) total += c->ws_size[i];
        if (total + bytes > c->ws_limit) return nullptr;
    }
    if (c->pins > 0) {          // a captured graph reads the current address
        c->pin_blocked = true;
        return nullptr;
    }
    if (c->ws[slot]) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->ws[slot]);
        c->ws[slot] = nullptr;
        c->ws_size[slot] = 0;
    }
    size_t sz = bytes < 256 ? 256 : bytes;
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    c->ws[slot] = p;
    c->ws_size[slot] = sz;
    return p;
}

extern "C" int xr_set_memory_limit(xr_ctx* c, int64_t bytes) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (bytes < 0) return xr_fail(c, XR_EINVAL, "negative memory limit");
    c->ws_limit = (size_t)bytes;
    return XR_OK;
}

extern "C" int xr_assume_hermitian_v(xr_ctx* c, int on) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->hermitian_v = on != 0;
    return XR_OK;
}

extern "C" int xr_trim(xr_ctx* c) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (c->pins > 0) return xr_fail(c, XR_EBUSY, "xr_trim: a captured graph holds the workspaces (xr_pin)");
    XR_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->ws.size(); i++) {
        if (c->ws[i]) XR_HIP(c, hipFree(c->ws[i]));
        c->ws[i] = nullptr;
        c->ws_size[i] = 0;
    }
    return XR_OK;
}

extern "C" int xr_pin(xr_ctx* c, int delta) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (c->pins + delta < 0) return xr_fail(c, XR_EINVAL, "xr_pin: more releases than pins");
    c->pins += delta;
    return XR_OK;
}

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
static hipEvent_t ev_get(xr_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

ProfScope::ProfScope(xr_ctx* c_, int kind, double flops, double bytes) : c(c_) {
    if (!c->prof || kind < 0 || kind >= 32 || !((c->prof_kinds >> kind) & 1u)) return;
    ProfRec r;
    r.kind = kind;
    r.flops = flops;
    r.bytes = bytes;
    r.a = ev_get(c);
    r.b = ev_get(c);
    if (!r.a || !r.b) return;
    (void)hipEventRecord(r.a, c->stream);
    c->recs.push_back(r);
    idx = (int)c->recs.size() - 1;
}

void ProfScope::shape(int64_t m, int64_t n, int64_t k, int64_t b, int ta, int tb) {
    if (idx < 0) return;
    ProfRec& r = c->recs[idx];
    r.m = m;
    r.n = n;
    r.k = k;
    r.batch = b;
    r.ta = ta;
    r.tb = tb;
}

ProfScope::~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(c->recs[idx].b, c->stream);
}

extern "C" int xr_profile(xr_ctx* c, int enable) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (enable) {
        XR_HIP(c, hipStreamSynchronize(c->stream));
        for (auto& r : c->recs) {
            c->ev_pool.push_back(r.a);
            c->ev_pool.push_back(r.b);
        }
        c->recs.clear();
    }
    c->prof = enable != 0;
    return XR_OK;
}

extern "C" int xr_profile_kinds(xr_ctx* c, uint32_t mask) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->prof_kinds = mask;
    return XR_OK;
}

extern "C" int xr_profile_stats(xr_ctx* c, int kind, int64_t* launches, double* total_ms,
                                double* flops, double* bytes) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    XR_HIP(c, hipStreamSynchronize(c->stream));
    int64_t n = 0;
    double ms = 0, fl = 0, by = 0;
    for (auto& r : c->recs) {
        if (r.kind != kind) continue;
        float t = 0;
        XR_HIP(c, hipEventElapsedTime(&t, r.a, r.b));
        n++;
        ms += t;
        fl += r.flops;
        by += r.bytes;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (flops) *flops = fl;
    if (bytes) *bytes = by;
    return XR_OK;
}

extern "C" int xr_profile_dump(xr_ctx* c, const char* path) {
    if (!c || !path) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    XR_HIP(c, hipStreamSynchronize(c->stream));
    FILE* f = fopen(path, "w");
    if (!f) return xr_fail(c, XR_EINVAL, std::string("cannot open ") + path);
    fprintf(f, "kind,m,n,k,batch,ta,tb,ms,flops,bytes\n");
    for (auto& r : c->recs) {
        float t = 0;
        if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) t = -1;
        fprintf(f, "%d,%lld,%lld,%lld,%lld,%d,%d,%.6f,%.6g,%.6g\n", r.kind, (long long)r.m,
                (long long)r.n, (long long)r.k, (long long)r.batch, r.ta, r.tb, t, r.flops, r.bytes);
    }
    fclose(f);
    return XR_OK;
}

// Plain fill kernel, used instead of hipMemsetAsync so the work can sit
// inside a captured HIP graph.
__global__ void k_fill(double* p, int64_t n, double v) {
    XR_GRID_STRIDE(i, n) p[i] = v;
}

int xr_fill(xr_ctx* c, double* p, int64_t n, double v) {
    if (n <= 0) return XR_OK;
    k_fill<<<xr_grid(n), 256, 0, c->stream>>>(p, n, v);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// ---------------------------------------------------------------------------
// configuration lists -> masks, with checks
// ---------------------------------------------------------------------------
// flags bit 0: a row is not strictly ascending / repeats an orbital (Pauli)
//       bit 1: orbital index out of range
//       bit 2: a core spin orbital is empty
//       bit 3: row i does not have lexical rank i
__device__ __forceinline__ void k_cfg_to_masks_elem(int64_t i, const int64_t* cfg, int64_t n_cfg, int n_elec, int n_orbs,
                               int n_core, uint64_t* masks, unsigned int* flags) {
    int dim = 2 * n_orbs;
    uint64_t m = 0;
    unsigned int f = 0;
    int64_t prev = -1;
    for (int e = 0; e < n_elec; e++) {
        int64_t p = cfg[i * n_elec + e];
        if (p < 0 || p >= dim) { f |= 2u; continue; }
        if (p <= prev) f |= 1u;
        prev = p;
        m |= 1ull << p;
    }
    uint64_t core = d_core_mask(n_orbs, n_core);
    if ((m & core) != core) f |= 4u;
    if (!f) {
        int64_t r = d_lex_rank(d_to_valence(m, n_orbs, n_core), dim - 2 * n_core);
        if (r != i) f |= 8u;
    }
    masks[i] = m;
    if (f) atomicOr(flags, f);
}

__global__ void k_cfg_to_masks(const int64_t* cfg, int64_t n_cfg, int n_elec, int n_orbs,
                               int n_core, uint64_t* masks, unsigned int* flags) {
    XR_GRID_STRIDE(i, n_cfg) k_cfg_to_masks_elem(i, cfg, n_cfg, n_elec, n_orbs, n_core, masks, flags);
}

extern "C" int xr_configs_to_masks(xr_ctx* c, const int64_t* d_cfg, int64_t n_cfg, int n_elec,
                                   int n_orbs, int n_core, uint64_t* d_masks) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n_orbs <= 0 || 2 * n_orbs > XR_MAX_SO || n_core < 0 || n_core > n_orbs)
        return xr_fail(c, XR_EUNSUPPORTED, "need 0 < 2*n_orbs <= 64 and 0 <= n_core <= n_orbs");
    if (n_elec < 2 * n_core || n_elec > 2 * n_orbs)
        return xr_fail(c, XR_EINVAL, "electron count incompatible with orbitals / core");
    int64_t expect = h_binom(2 * (n_orbs - n_core), n_elec - 2 * n_core);
    if (n_cfg != expect)
        return xr_fail(c, XR_EINVAL_CONFIG_ORDER,
                       "config list is not the complete determinant space (expected " +
                           std::to_string(expect) + " rows, got " + std::to_string(n_cfg) + ")");
    unsigned int* flags = (unsigned int*)xr_workspace(c, 15, sizeof(unsigned int));
    if (!flags) return xr_fail(c, XR_ENOMEM, "workspace");
    XR_HIP(c, hipMemsetAsync(flags, 0, sizeof(unsigned int), c->stream));
    k_cfg_to_masks<<<xr_grid(n_cfg), 256, 0, c->stream>>>(d_cfg, n_cfg, n_elec, n_orbs,
                                                                 n_core, d_masks, flags);
    XR_LAUNCH_CHECK(c);
    unsigned int hf = 0;
    XR_HIP(c, hipMemcpyAsync(&hf, flags, sizeof(hf), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    if (hf & 1u) return xr_fail(c, XR_EPAULI, "a configuration repeats a spin orbital or is not ascending");
    if (hf & 2u) return xr_fail(c, XR_EINVAL, "spin-orbital index out of range");
    if (hf & 4u) return xr_fail(c, XR_EINVAL_CONFIG_ORDER, "a configuration has an empty core orbital");
    if (hf & 8u) return xr_fail(c, XR_EINVAL_CONFIG_ORDER, "configurations are not in lexical order");
    return XR_OK;
}

__device__ __forceinline__ void k_rank_elem(int64_t i, const uint64_t* masks, int64_t n, int n_orbs, int n_core, int64_t* out) {
    out[i] = d_lex_rank(d_to_valence(masks[i], n_orbs, n_core), 2 * (n_orbs - n_core));
}

__global__ void k_rank(const uint64_t* masks, int64_t n, int n_orbs, int n_core, int64_t* out) {
    XR_GRID_STRIDE(i, n) k_rank_elem(i, masks, n, n_orbs, n_core, out);
}

extern "C" int xr_config_rank(xr_ctx* c, const uint64_t* d_masks, int64_t n, int n_orbs,
                              int n_core, int64_t* d_rank) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n_orbs <= 0 || 2 * n_orbs > XR_MAX_SO) return xr_fail(c, XR_EUNSUPPORTED, "dim > 64");
    if (n <= 0) return XR_OK;
    k_rank<<<xr_grid(n), 256, 0, c->stream>>>(d_masks, n, n_orbs, n_core, d_rank);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// ---------------------------------------------------------------------------
// context lifecycle
// ---------------------------------------------------------------------------
extern "C" int xr_create(int device, void* stream, xr_ctx** out) {
    if (!out) return XR_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return XR_EHIP;
    if (hipSetDevice(device) != hipSuccess) return XR_EHIP;
    xr_ctx* c = new xr_ctx();
    c->device = device;
    c->stream = (hipStream_t)stream;
    static uint64_t tbl[65 * 65];
    for (int n = 0; n <= 64; n++)
        for (int k = 0; k <= 64; k++) {
            // Pascal's rule in 64-bit; every entry up to C(64,32) fits
            if (k == 0 || k == n) tbl[n * 65 + k] = 1;
            else if (k > n) tbl[n * 65 + k] = 0;
            else tbl[n * 65 + k] = tbl[(n - 1) * 65 + k - 1] + tbl[(n - 1) * 65 + k];
        }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_binom), tbl, sizeof(tbl)) != hipSuccess) {
        delete c;
        return XR_EHIP;
    }
    if (xr_noalloc_init(c) != XR_OK) {     // the device address of the scratch array
        delete c;
        return XR_EHIP;
    }
    *out = c;
    return XR_OK;
}

extern "C" int xr_destroy(xr_ctx* c) {
    if (!c) return XR_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void* p : c->ws)
        if (p) (void)hipFree(p);
    for (auto& kv : c->kspace) (void)hipFree(kv.second.first);
    for (auto& kv : c->sidetab) (void)hipFree(kv.second.first);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->pair_copied) (void)hipEventDestroy(c->pair_copied);
    if (c->pair_host) (void)hipHostFree(c->pair_host);
    delete c;
    return XR_OK;
}

extern "C" int xr_set_stream(xr_ctx* c, void* stream) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->stream = (hipStream_t)stream;
    return XR_OK;
}

extern "C" const char* xr_last_error(xr_ctx* c) {
    if (!c) return "null context";
    return c->err.c_str();
}

extern "C" int xr_synchronize(xr_ctx* c) {
    if (!c) return XR_EINVAL;
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}
