// Transition densities as gather-GEMMs on the matrix cores.
//
// The reference walks every (ket determinant, operator-index tuple) and
// scatters a rank-1 update into all state pairs (general-XRCC/
// density_tensors.c:142-556).  Here an operator string is split as
// L * R with L pure creators (possibly empty) and R pure annihilators or
// pure creators, and a resolution of the identity over the complete
// determinant space K between them is inserted:
//
//   <I| L_l R_t |J> = sum_K <I|L_l|K> <K|R_t|J>
//                   = sum_K X[K, l, I] * Y[K, t, J]
//
//   X[K, l, I] = sum_P <K| L_l^+ |P> zbra[I, P]   (one P per (K, l))
//   Y[K, t, J] = sum_Q <K| R_t |Q> zket[J, Q]     (one Q per (K, t))
//
// l and t run over sorted index tuples only (operators of one kind
// anticommute), so X and Y are signed gathers of coefficient rows driven by
// integer excitation tables, and the contraction over K is one dense FP64
// GEMM.  An expand pass writes the full antisymmetric dim^k layout.  When L
// is empty, K is the bra space itself and X = zbra^T.
#include "xr_common.h"
#include "xr_gemm.h"
#include "xr_density.h"

// ---------------------------------------------------------------------------
// operator plans
// ---------------------------------------------------------------------------
int xr_plan_op(const char* ops, OpPlan* p) {
    struct Row { const char* s; OpPlan p; };
    // side: {k, is_c, slots}; slots name the tuple entry each operator of
    // the side string uses, leftmost operator first.
    static const Row rows[] = {
        {"a",    {1, 0, {0, 0, {0}}, {1, 0, {0}}}},
        {"c",    {1, 1, {1, 0, {0}}, {0, 0, {0}}}},
        {"aa",   {2, 0, {0, 0, {0}}, {2, 0, {0, 1}}}},
        {"cc",   {2, 1, {2, 0, {1, 0}}, {0, 0, {0}}}},
        {"ca",   {2, 1, {1, 0, {0}}, {1, 0, {0}}}},
        {"caa",  {3, 1, {1, 0, {0}}, {2, 0, {0, 1}}}},
        {"cca",  {3, 1, {2, 0, {1, 0}}, {1, 0, {0}}}},
        {"ccaa", {4, 1, {2, 0, {1, 0}}, {2, 0, {0, 1}}}},
        {"caaa", {4, 1, {1, 0, {0}}, {3, 0, {0, 1, 2}}}},
        {"ccca", {4, 1, {3, 0, {2, 1, 0}}, {1, 0, {0}}}},
        // the 5-operator strings of XR order 1 (Be-states/densities.py:83-88;
        // Be-states/_attic_/deprecated/density_tensors.c:690-836)
        {"ccaaa", {5, 1, {2, 0, {1, 0}}, {3, 0, {0, 1, 2}}}},
        {"cccaa", {5, 1, {3, 0, {2, 1, 0}}, {2, 0, {0, 1}}}},
    };
    if (!ops) return 0;
    for (const Row& r : rows) {
        if (strcmp(ops, r.s) == 0) {
            *p = r.p;
            int d = 0;
            for (const char* q = ops; *q; q++) d += (*q == 'c') ? 1 : -1;
            p->delta = d;
            return 1;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t d_from_valence(uint64_t v, int n_orbs, int n_core) {
    if (n_core == 0) return v;
    int nv = n_orbs - n_core;
    uint64_t m = (nv >= 64) ? ~0ull : ((1ull << nv) - 1);
    uint64_t a = (v & m) << n_core;
    uint64_t b = ((v >> nv) & m) << (n_orbs + n_core);
    return a | b | d_core_mask(n_orbs, n_core);
}

__device__ __forceinline__ void k_unrank_space_elem(int64_t i, uint64_t* masks, int64_t n, int n_orbs, int n_core, int n_elec) {
    int nval = 2 * (n_orbs - n_core);
    masks[i] = d_from_valence(d_lex_unrank(i, nval, n_elec - 2 * n_core), n_orbs, n_core);
}

__global__ void k_unrank_space(uint64_t* masks, int64_t n, int n_orbs, int n_core, int n_elec) {
    XR_GRID_STRIDE(i, n) k_unrank_space_elem(i, masks, n, n_orbs, n_core, n_elec);
}

// One entry per (K, t): the source determinant the side string maps onto K,
// as sign * (valence rank + 1), or 0 when the string annihilates it.
__device__ __forceinline__ void k_side_table_elem(int64_t i, SidePlan sd, const uint64_t* __restrict__ Km, int64_t nK, int dim,
                             int n_orbs, int n_core, int64_t nT, int32_t* __restrict__ tab) {
    int64_t K = i / nT, t = i - K * nT;
    uint64_t km = Km[K];
    uint64_t tm = d_lex_unrank(t, dim, sd.k);
    int orb[3] = {0, 0, 0};
    {
        uint64_t x = tm;
        for (int j = 0; j < sd.k; j++) {
            orb[j] = __ffsll((unsigned long long)x) - 1;
            x &= x - 1;
        }
    }
    uint64_t src;
    if (!sd.is_c) {
        if (km & tm) { tab[i] = 0; return; }
        src = km | tm;
    } else {
        if ((km & tm) != tm) { tab[i] = 0; return; }
        src = km & ~tm;
    }
    uint64_t core = d_core_mask(n_orbs, n_core);
    if ((src & core) != core) { tab[i] = 0; return; }
    uint64_t cur = src;
    int sg = 1;
    for (int j = sd.k - 1; j >= 0; j--) {
        int o = orb[sd.slot[j]];
        sg *= d_sign_above(cur, o);
        cur ^= 1ull << o;
    }
    int64_t r = d_lex_rank(d_to_valence(src, n_orbs, n_core), dim - 2 * n_core);
    tab[i] = sg * (int32_t)(r + 1);
}

__global__ void k_side_table(SidePlan sd, const uint64_t* __restrict__ Km, int64_t nK, int dim,
                             int n_orbs, int n_core, int64_t nT, int32_t* __restrict__ tab) {
    XR_GRID_STRIDE(i, nK * nT) k_side_table_elem(i, sd, Km, nK, dim, n_orbs, n_core, nT, tab);
}

// zT[c * nS + s] = z[s * nC + c]
__global__ void k_transpose(const double* __restrict__ z, int64_t nS, int64_t nC,
                            double* __restrict__ zT) {
    __shared__ double tile[32][33];
    int64_t c0 = (int64_t)blockIdx.x * 32, s0 = (int64_t)blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        int64_t s = s0 + j, c = c0 + tx;
        if (s < nS && c < nC) tile[j][tx] = z[s * nC + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        int64_t c = c0 + j, s = s0 + tx;
        if (s < nS && c < nC) zT[c * nS + s] = tile[tx][j];
    }
}

// Signed row gather: row r = K * nT + t of the output is sign * row |e| - 1
// of zT (nS states) for e = tab[r], zero for e = 0.  layout 0: out[K][t][s];
// layout 1: out[t][K][s].  A wave copies 64 / SP rows at a time (SP = nS
// rounded up to a power of two, at most 64; lanes (row, s)), so the
// per-row index work is done once per row, not per element, and every row
// is one contiguous read and one contiguous write.
template <int SP>
__global__ __launch_bounds__(256) void k_gather_rows(const int32_t* __restrict__ tab, int64_t nK, int64_t nT,
                                                     const double* __restrict__ zT, int64_t nS,
                                                     double* __restrict__ out, int layout) {
    constexpr int RPW = 64 / SP;                 // rows per wave and step
    const int lane = threadIdx.x & 63;
    const int sub = lane / SP, s0 = lane % SP;
    const int64_t rows = nK * nT;
    const int64_t wave = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW;
    const int64_t step = (int64_t)gridDim.x * 4 * RPW;
    for (int64_t r0 = wave; r0 < rows; r0 += step) {
        const int64_t r = r0 + sub;
        if (r >= rows) continue;
        const int32_t e = tab[r];
        const int64_t K = r / nT, t = r - K * nT;
        double* dst = out + (layout == 0 ? r : t * nK + K) * nS;
        if (e == 0) {
            for (int64_t s = s0; s < nS; s += SP) dst[s] = 0.0;
        } else {
            const double sg = e > 0 ? 1.0 : -1.0;
            const double* src = zT + (int64_t)((e > 0 ? e : -e) - 1) * nS;
            for (int64_t s = s0; s < nS; s += SP) dst[s] = sg * src[s];
        }
    }
}


// For each flat index (d_0 ... d_{k-1}) in [0, dim^k): the lexical rank of
// the sorted tuple and the parity of the sort, as sign * (rank + 1); 0 if an
// index repeats.
__device__ __forceinline__ void k_index_table_elem(int64_t i, int k, int dim, int64_t n, int32_t* tab) {
    int d[3];
    int64_t x = i;
    for (int j = k - 1; j >= 0; j--) {
        d[j] = (int)(x % dim);
        x /= dim;
    }
    int inv = 0;
    uint64_t m = 0;
    for (int a = 0; a < k; a++) {
        for (int b = a + 1; b < k; b++) {
            if (d[a] == d[b]) { tab[i] = 0; return; }
            if (d[a] > d[b]) inv++;
        }
        m |= 1ull << d[a];
    }
    int32_t r = (int32_t)d_lex_rank(m, dim);
    tab[i] = ((inv & 1) ? -1 : 1) * (r + 1);
}

__global__ void k_index_table(int k, int dim, int64_t n, int32_t* tab) {
    XR_GRID_STRIDE(i, n) k_index_table_elem(i, k, dim, n, tab);
}

// out[((I*nJ + J)*nL + iL)*nR + iR] from the compressed GEMM result
// Cc[(l*nI + I) * ldc + t*nJ + J], through the signed index tables
// tab[i] = sign * (sorted-tuple rank + 1), 0 for a repeated index.
// blockIdx.x walks the (I, J) planes (J fastest, remapped so that each XCD
// takes a contiguous run of planes) and blockIdx.y 1024-element chunks of a
// plane: the blocks in flight on one XCD write contiguous chunks of
// neighbouring planes and read the same Cc rows for consecutive J, so one
// cache line of Cc serves 8 planes in that XCD's L2.  Plane offsets are 32-bit.
__global__ __launch_bounds__(256) void k_expand_planes(
    const double* __restrict__ Cc, int64_t ldc, int64_t nI, int64_t nJ,
    const int32_t* __restrict__ tabL, uint32_t nL, const int32_t* __restrict__ tabR, uint32_t nR,
    double* __restrict__ out) {
    const uint32_t plane = nL * nR;
    const int64_t n_planes = nI * nJ;
    const int64_t lstride = nI * ldc;
    for (int64_t b = blockIdx.x; b < n_planes; b += gridDim.x) {
        const int64_t pl = (b - blockIdx.x) + xcd_remap(blockIdx.x, gridDim.x);
        if (pl >= n_planes) continue;
        const int64_t I = pl / nJ, J = pl - I * nJ;
        const double* src = Cc + I * ldc + J;
        double* dst = out + pl * (int64_t)plane;
        for (uint32_t base = blockIdx.y * 1024u; base < plane; base += gridDim.y * 1024u) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t e = base + k * 256u + threadIdx.x;
                if (e >= plane) break;
                const uint32_t iL = e / nR, iR = e - iL * nR;
                const int32_t eL = tabL ? tabL[iL] : 1;
                const int32_t eR = tabR ? tabR[iR] : 1;
                double v = 0.0;
                if (eL != 0 && eR != 0) {
                    const int64_t l = (eL > 0 ? eL : -eL) - 1, t = (eR > 0 ? eR : -eR) - 1;
                    v = src[l * lstride + t * nJ];
                    if ((eL < 0) != (eR < 0)) v = -v;
                }
                dst[e] = v;
            }
        }
    }
}

// Veff[u, w] for sorted pairs u = (p<q), w = (a<b):
//   V[p,q,b,a] - V[q,p,b,a] - V[p,q,a,b] + V[q,p,a,b]
__device__ __forceinline__ void k_veff_elem(int64_t i, const double* __restrict__ V, int dim, int64_t nU,
                       double* __restrict__ Ve) {
    int64_t u = i / nU, w = i - u * nU;
    uint64_t mu = d_lex_unrank(u, dim, 2), mw = d_lex_unrank(w, dim, 2);
    int64_t p = __ffsll((unsigned long long)mu) - 1;
    int64_t q = 63 - __clzll((long long)mu);
    int64_t a = __ffsll((unsigned long long)mw) - 1;
    int64_t b = 63 - __clzll((long long)mw);
    int64_t D = dim;
    auto at = [&](int64_t x0, int64_t x1, int64_t x2, int64_t x3) {
        return V[((x0 * D + x1) * D + x2) * D + x3];
    };
    Ve[i] = at(p, q, b, a) - at(q, p, b, a) - at(p, q, a, b) + at(q, p, a, b);
}

__global__ void k_veff(const double* __restrict__ V, int dim, int64_t nU,
                       double* __restrict__ Ve) {
    XR_GRID_STRIDE(i, nU * nU) k_veff_elem(i, V, dim, nU, Ve);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
int xr_space_masks(xr_ctx* c, int n_orbs, int n_core, int n_elec, const uint64_t** masks,
                   int64_t* n) {
    auto key = std::make_pair(n_orbs * 1024 + n_core, n_elec);
    auto it = c->kspace.find(key);
    if (it != c->kspace.end()) {
        *masks = it->second.first;
        *n = it->second.second;
        return XR_OK;
    }
    int64_t cnt = h_binom(2 * (n_orbs - n_core), n_elec - 2 * n_core);
    if (cnt <= 0) return xr_fail(c, XR_EINVAL, "empty determinant space");
    uint64_t* p = nullptr;
    XR_HIP(c, hipMalloc(&p, cnt * sizeof(uint64_t)));
    k_unrank_space<<<xr_grid(cnt), 256, 0, c->stream>>>(p, cnt, n_orbs, n_core, n_elec);
    XR_LAUNCH_CHECK(c);
    c->kspace[key] = std::make_pair(p, cnt);
    *masks = p;
    *n = cnt;
    return XR_OK;
}

int xr_transpose(xr_ctx* c, const double* z, int64_t nS, int64_t nC, double* zT) {
    if (nS <= 0 || nC <= 0) return XR_OK;
    dim3 grid((unsigned)ceil_div(nC, 32), (unsigned)ceil_div(nS, 32));
    k_transpose<<<grid, dim3(32, 8), 0, c->stream>>>(z, nS, nC, zT);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// The side table of (sd, K space) depends on the determinant spaces only:
// built on first use and kept for the context's lifetime (*tab; nullptr
// for an empty table).
int xr_side_table(xr_ctx* c, const SidePlan& sd, const uint64_t* Km, int64_t nK, int n_orbs,
                  int n_core, int32_t** tab) {
    int64_t nT = h_binom(2 * n_orbs, sd.k);
    int64_t n = nK * nT;
    *tab = nullptr;
    if (n <= 0) return XR_OK;
    const std::array<int64_t, 12> key{sd.k, sd.is_c, sd.slot[0], sd.slot[1], sd.slot[2],
                                      (int64_t)(uintptr_t)Km, nK, n_orbs, n_core, -1, 0, 0};
    auto it = c->sidetab.find(key);
    if (it != c->sidetab.end()) {
        *tab = it->second.first;
        return XR_OK;
    }
    int32_t* p = nullptr;
    XR_HIP(c, hipMalloc(&p, n * sizeof(int32_t)));
    c->sidetab[key] = std::make_pair(p, n);
    ProfScope prof(c, XR_K_TABLE, 0.0, 4.0 * n);
    k_side_table<<<xr_grid(n), 256, 0, c->stream>>>(sd, Km, nK, 2 * n_orbs, n_orbs, n_core, nT, p);
    XR_LAUNCH_CHECK(c);
    *tab = p;
    return XR_OK;
}

int xr_gather(xr_ctx* c, const int32_t* tab, int64_t nK, int64_t nT, const double* zT,
              int64_t nS, double* out, int layout) {
    int64_t n = nK * nT * nS;
    if (n <= 0) return XR_OK;
    ProfScope prof(c, XR_K_GATHER, 0.0, 8.0 * n + 4.0 * nK * nT);
    const int sp = nS <= 1 ? 1 : nS <= 2 ? 2 : nS <= 4 ? 4 : nS <= 8 ? 8 : nS <= 16 ? 16 : nS <= 32 ? 32 : 64;
    const int64_t waves = ceil_div(nK * nT, 64 / sp);
    const unsigned grid = xr_grid(waves * 64);
    switch (sp) {
    case 1: k_gather_rows<1><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    case 2: k_gather_rows<2><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    case 4: k_gather_rows<4><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    case 8: k_gather_rows<8><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    case 16: k_gather_rows<16><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    case 32: k_gather_rows<32><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    default: k_gather_rows<64><<<grid, 256, 0, c->stream>>>(tab, nK, nT, zT, nS, out, layout); break;
    }
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

int xr_index_table(xr_ctx* c, int k, int dim, int32_t* tab) {
    int64_t n = 1;
    for (int j = 0; j < k; j++) n *= dim;
    k_index_table<<<xr_grid(n), 256, 0, c->stream>>>(k, dim, n, tab);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

static int check_spaces(xr_ctx* c, const OpPlan& pl, int n_orbs, int n_core, int n_elec_ket,
                        int64_t nP, int64_t nQ, int* n_elec_bra) {
    if (n_orbs <= 0 || 2 * n_orbs > XR_MAX_SO) return xr_fail(c, XR_EUNSUPPORTED, "need 0 < 2*n_orbs <= 64");
    if (n_core < 0 || n_core > n_orbs) return xr_fail(c, XR_EINVAL, "bad n_core");
    int nb = n_elec_ket + pl.delta;
    if (n_elec_ket < 2 * n_core || n_elec_ket > 2 * n_orbs || nb < 2 * n_core || nb > 2 * n_orbs)
        return xr_fail(c, XR_EINVAL, "electron counts incompatible with orbitals / core");
    int nval = 2 * (n_orbs - n_core);
    if (nQ != h_binom(nval, n_elec_ket - 2 * n_core))
        return xr_fail(c, XR_ESHAPE, "ket coefficient columns != size of the ket determinant space");
    if (nP != h_binom(nval, nb - 2 * n_core))
        return xr_fail(c, XR_ESHAPE, "bra coefficient columns != size of the bra determinant space");
    if (nP >= 0x7fffffff || nQ >= 0x7fffffff) return xr_fail(c, XR_EUNSUPPORTED, "determinant space >= 2^31");
    *n_elec_bra = nb;
    return XR_OK;
}

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
enum { WS_ZTB = 0, WS_ZTK, WS_TABL, WS_TABR, WS_X, WS_Y, WS_C, WS_IDXL, WS_IDXR, WS_VE, WS_W };
enum { WS_GC = 40, WS_XB, WS_YB, WS_ZBT, WS_WB, WS_VEB, WS_GS };

// The two gathered factors of a density: X (nK x nLt*nI) from the bra
// coefficients and Y (nK x nT*nJ) from the ket coefficients, over the K
// space that sits between L and R: the bra space when there is no L, the
// ket space when there is no R (then X or Y is just the transposed
// coefficients), else every determinant with n_ket - |R| electrons.
struct Factors {
    const uint64_t* Km = nullptr;
    int64_t nK = 0, nLt = 1, nT = 1;
    int kspace_core = 0;
    const double* X = nullptr;
    const double* Y = nullptr;
};

static int prepare_factors(xr_ctx* c, const OpPlan& pl, int n_orbs, int n_core, int n_elec_ket,
                           int nb, int64_t nI, int64_t nP, const double* zbra, int64_t nJ,
                           int64_t nQ, const double* zket, Factors* f) {
    const int dim = 2 * n_orbs;
    int rc;
    double* zTb = (double*)xr_workspace(c, WS_ZTB, nP * nI * sizeof(double));
    double* zTk = (double*)xr_workspace(c, WS_ZTK, nQ * nJ * sizeof(double));
    if (!zTb || !zTk) return xr_fail(c, XR_ENOMEM, "workspace (coefficients)");
    if ((rc = xr_transpose(c, zbra, nI, nP, zTb))) return rc;
    if ((rc = xr_transpose(c, zket, nJ, nQ, zTk))) return rc;
    f->nT = pl.R.k ? h_binom(dim, pl.R.k) : 1;
    f->nLt = pl.L.k ? h_binom(dim, pl.L.k) : 1;
    if (!pl.L.k) {
        if ((rc = xr_space_masks(c, n_orbs, n_core, nb, &f->Km, &f->nK))) return rc;
        f->X = zTb;
        f->kspace_core = n_core;
    } else if (!pl.R.k) {
        if ((rc = xr_space_masks(c, n_orbs, n_core, n_elec_ket, &f->Km, &f->nK))) return rc;
        f->Y = zTk;
        f->kspace_core = n_core;
    } else {
        if ((rc = xr_space_masks(c, n_orbs, 0, n_elec_ket - pl.R.k, &f->Km, &f->nK))) return rc;
        f->kspace_core = 0;
    }
    int32_t* tabL = nullptr;
    if (pl.L.k) {
        double* Xb = (double*)xr_workspace(c, WS_X, f->nK * f->nLt * nI * sizeof(double));
        if (!Xb) return xr_fail(c, XR_ENOMEM, "workspace (bra gather)");
        if ((rc = xr_side_table(c, pl.L, f->Km, f->nK, n_orbs, n_core, &tabL))) return rc;
        if ((rc = xr_gather(c, tabL, f->nK, f->nLt, zTb, nI, Xb, 0))) return rc;
        f->X = Xb;
    }
    if (pl.R.k) {
        int32_t* tabR = nullptr;
        if ((rc = xr_side_table(c, pl.R, f->Km, f->nK, n_orbs, n_core, &tabR))) return rc;
        if (pl.L.k && tabR == tabL && zbra == zket && nI == nJ && nP == nQ) {
            f->Y = f->X;                    // a block with itself through one table: the same factor
        } else {
            double* Yb = (double*)xr_workspace(c, WS_Y, f->nK * f->nT * nJ * sizeof(double));
            if (!Yb) return xr_fail(c, XR_ENOMEM, "workspace (ket gather)");
            if ((rc = xr_gather(c, tabR, f->nK, f->nT, zTk, nJ, Yb, 0))) return rc;
            f->Y = Yb;
        }
    }
    return XR_OK;
}

int xr_density_impl(xr_ctx* c, const char* ops, int n_orbs, int n_core, int n_elec_ket,
                    int64_t nI, int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                    const double* zket, double* out) {
    OpPlan pl;
    if (!xr_plan_op(ops, &pl)) return xr_fail(c, XR_EUNSUPPORTED, std::string("operator string not supported: ") + (ops ? ops : "(null)"));
    int nb = 0;
    int rc = check_spaces(c, pl, n_orbs, n_core, n_elec_ket, nP, nQ, &nb);
    if (rc) return rc;
    const int dim = 2 * n_orbs;
    if (nI <= 0 || nJ <= 0) return XR_OK;
    int64_t out_size = nI * nJ;
    for (int j = 0; j < pl.k; j++) out_size *= dim;
    if (pl.L.k && pl.R.k && (n_elec_ket - pl.R.k < 0))  // R annihilates more electrons than the ket has
        return xr_fill(c, out, out_size, 0.0);
    Factors f;
    if ((rc = prepare_factors(c, pl, n_orbs, n_core, n_elec_ket, nb, nI, nP, zbra, nJ, nQ, zket, &f)))
        return rc;
    const int64_t Mg = f.nLt * nI, Ng = f.nT * nJ;
    double* Cc = (double*)xr_workspace(c, WS_C, Mg * Ng * sizeof(double));
    if (!Cc) return xr_fail(c, XR_ENOMEM, "workspace (compressed density)");
    if (f.X == f.Y && Mg == Ng) {
        // a block with itself through one factor: Cc = X^T X is symmetric
        if ((rc = xr_dgemm_syrk_tn(c, Mg, f.nK, 1.0, f.X, Mg, Cc, Ng))) return rc;
    } else if ((rc = xr_dgemm_launch(c, 1, 0, Mg, Ng, f.nK, 1.0, f.X, Mg, 0, f.Y, Ng, 0, 0.0, Cc, Ng, 0, 1))) {
        return rc;
    }
    int64_t nIL = 1, nIR = 1;
    for (int j = 0; j < pl.L.k; j++) nIL *= dim;
    for (int j = 0; j < pl.R.k; j++) nIR *= dim;
    int32_t* idxL = pl.L.k ? (int32_t*)xr_workspace(c, WS_IDXL, nIL * sizeof(int32_t)) : nullptr;
    int32_t* idxR = pl.R.k ? (int32_t*)xr_workspace(c, WS_IDXR, nIR * sizeof(int32_t)) : nullptr;
    if ((pl.L.k && !idxL) || (pl.R.k && !idxR)) return xr_fail(c, XR_ENOMEM, "workspace (index tables)");
    if (pl.L.k && (rc = xr_index_table(c, pl.L.k, dim, idxL))) return rc;
    if (pl.R.k && (rc = xr_index_table(c, pl.R.k, dim, idxR))) return rc;
    int64_t total = nI * nJ * nIL * nIR;
    ProfScope prof(c, XR_K_EXPAND, 0.0, 8.0 * total);
    if (nIL * nIR > 0xffffffffLL - 65536LL * 1024) return xr_fail(c, XR_EUNSUPPORTED, "density plane too large");
    const int64_t planes = nI * nJ;
    const int64_t chunks = ceil_div(nIL * nIR, 1024);
    dim3 grid((unsigned)(planes < (1LL << 30) ? planes : (1LL << 30)),
              (unsigned)(chunks < 65535 ? chunks : 65535));
    k_expand_planes<<<grid, 256, 0, c->stream>>>(Cc, Ng, nI, nJ, idxL, (uint32_t)nIL, idxR,
                                                 (uint32_t)nIR, out);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

extern "C" int xr_density(xr_ctx* c, const char* ops, int n_orbs, int n_core, int n_elec_ket,
                          int64_t nI, int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                          const double* zket, double* out) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    return xr_density_impl(c, ops, n_orbs, n_core, n_elec_ket, nI, nP, zbra, nJ, nQ, zket, out);
}

// Wkeep: caller's buffer for W = Veff Y (nU * nK * nJ doubles, w_len), which
// the adjoint can take instead of recomputing it; nullptr: a workspace.
static int monomer_2e_impl(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI, int64_t nP,
                           const double* zbra, int64_t nJ, int64_t nQ, const double* zket,
                           const double* V, double* out, double* Wkeep, int64_t w_len) {
    OpPlan pl;
    xr_plan_op("ccaa", &pl);
    int nb = 0;
    int rc = check_spaces(c, pl, n_orbs, n_core, n_elec, nP, nQ, &nb);
    if (rc) return rc;
    if (nI <= 0 || nJ <= 0) return XR_OK;
    const int dim = 2 * n_orbs;
    if (n_elec < 2) return xr_fill(c, out, nI * nJ, 0.0);
    double* zTb = (double*)xr_workspace(c, WS_ZTB, nP * nI * sizeof(double));
    double* zTk = (double*)xr_workspace(c, WS_ZTK, nQ * nJ * sizeof(double));
    if (!zTb || !zTk) return xr_fail(c, XR_ENOMEM, "workspace (coefficients)");
    if ((rc = xr_transpose(c, zbra, nI, nP, zTb))) return rc;
    if ((rc = xr_transpose(c, zket, nJ, nQ, zTk))) return rc;
    const uint64_t* Km = nullptr;
    int64_t nK = 0;
    if ((rc = xr_space_masks(c, n_orbs, 0, n_elec - 2, &Km, &nK))) return rc;
    const int64_t nU = h_binom(dim, 2);
    int32_t* tabL = nullptr;
    int32_t* tabR = nullptr;
    double* X = (double*)xr_workspace(c, WS_X, nU * nK * nI * sizeof(double));
    double* Y = (double*)xr_workspace(c, WS_Y, nU * nK * nJ * sizeof(double));
    if (Wkeep && w_len != nU * nK * nJ) return xr_fail(c, XR_ESHAPE, "sigma factor buffer size");
    double* W = Wkeep ? Wkeep : (double*)xr_workspace(c, WS_W, nU * nK * nJ * sizeof(double));
    double* Ve = (double*)xr_workspace(c, WS_VE, nU * nU * sizeof(double));
    if (!X || !Y || !W || !Ve) return xr_fail(c, XR_ENOMEM, "workspace (sigma)");
    if ((rc = xr_side_table(c, pl.L, Km, nK, n_orbs, n_core, &tabL))) return rc;
    if ((rc = xr_side_table(c, pl.R, Km, nK, n_orbs, n_core, &tabR))) return rc;
    // a block with itself: the two sides' tables differ only by the order of
    // the two annihilators, so X = sg Y and one gather serves both
    const double sg = (pl.L.slot[0] == pl.R.slot[0]) ? 1.0 : -1.0;
    const bool same = zbra == zket && nI == nJ && nP == nQ && pl.L.k == 2 && pl.R.k == 2 &&
                      pl.L.is_c == pl.R.is_c && pl.L.slot[0] != pl.L.slot[1] && pl.R.slot[0] != pl.R.slot[1];
    if (!same && (rc = xr_gather(c, tabL, nK, nU, zTb, nI, X, 1))) return rc;   // X[u][K][I]
    if ((rc = xr_gather(c, tabR, nK, nU, zTk, nJ, Y, 1))) return rc;   // Y[w][K][J]
    k_veff<<<xr_grid(nU * nU), 256, 0, c->stream>>>(V, dim, nU, Ve);
    XR_LAUNCH_CHECK(c);
    // W[u][K*nJ + J] = sum_w Ve[u,w] Y[w][K*nJ + J]
    if ((rc = xr_gemm_2e(c, 0, 0, nU, nK * nJ, nU, 1.0, Ve, nU, 0, Y, nK * nJ, 0, 0.0, W,
                              nK * nJ, 0, 1)))
        return rc;
    // E[I,J] = sum_{(u,K)} X[(u,K), I] W[(u,K), J]
    if (same) return xr_gemm_2e(c, 1, 0, nI, nJ, nU * nK, sg, Y, nI, 0, W, nJ, 0, 0.0, out, nJ, 0, 1);
    return xr_gemm_2e(c, 1, 0, nI, nJ, nU * nK, 1.0, X, nI, 0, W, nJ, 0, 0.0, out, nJ, 0, 1);
}

extern "C" int xr_monomer_2e(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI, int64_t nP,
                             const double* zbra, int64_t nJ, int64_t nQ, const double* zket,
                             const double* V, double* out) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    return monomer_2e_impl(c, n_orbs, n_core, n_elec, nI, nP, zbra, nJ, nQ, zket, V, out, nullptr, 0);
}

extern "C" int xr_monomer_2e_keep(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI, int64_t nP,
                                  const double* zbra, int64_t nJ, int64_t nQ, const double* zket,
                                  const double* V, double* out, double* W, int64_t w_len) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!W) return xr_fail(c, XR_EINVAL, "null sigma factor buffer");
    return monomer_2e_impl(c, n_orbs, n_core, n_elec, nI, nP, zbra, nJ, nQ, zket, V, out, W, w_len);
}

// ---------------------------------------------------------------------------
// adjoints
// ---------------------------------------------------------------------------
// Orbitals of the lexicographic k-tuple of rank r (ascending).
__device__ __forceinline__ void tuple_orbs(int64_t r, int dim, int k, int* o) {
    if (!k) return;
    uint64_t m = d_lex_unrank(r, dim, k);
    for (int j = 0; j < k; j++) { o[j] = __ffsll((unsigned long long)m) - 1; m &= m - 1; }
}

// Sum over the index orderings of the sorted tuples lo (kL) and ro (kR) of
// sign * plane[iL * nIR + iR].
__device__ __forceinline__ double compress_sum(const double* __restrict__ plane, const int* lo, const int* ro,
                                               int kL, int kR, int dim, int64_t nIR) {
    const int P3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int S3[6] = {1, -1, -1, 1, 1, -1};
    const int P2[2][3] = {{0, 1, 0}, {1, 0, 0}};
    const int S2[2] = {1, -1};
    const int nPL = kL == 3 ? 6 : (kL == 2 ? 2 : 1);
    const int nPR = kR == 3 ? 6 : (kR == 2 ? 2 : 1);
    double acc = 0.0;
    for (int a = 0; a < nPL; a++) {
        int64_t iL = 0;
        for (int j = 0; j < kL; j++) iL = iL * dim + lo[kL == 3 ? P3[a][j] : (kL == 2 ? P2[a][j] : 0)];
        const int sL = kL == 3 ? S3[a] : (kL == 2 ? S2[a] : 1);
        for (int b = 0; b < nPR; b++) {
            int64_t iR = 0;
            for (int j = 0; j < kR; j++) iR = iR * dim + ro[kR == 3 ? P3[b][j] : (kR == 2 ? P2[b][j] : 0)];
            const int sR = kR == 3 ? S3[b] : (kR == 2 ? S2[b] : 1);
            acc += (double)(sL * sR) * plane[iL * nIR + iR];
        }
    }
    return acc;
}

// Transpose of k_expand: Gc[(l*nI + I) * ldc + t*nJ + J] = sum over the index
// orderings of l and t of sign * Rbar[I, J, iL, iR].  Rbar is read along t
// inside one (I, J) plane and Gc written along J: a workgroup takes 32
// consecutive J by 32 consecutive t for one (l, I) and turns the tile
// through LDS.  blockIdx.x = (I, J tile), blockIdx.y = l, blockIdx.z = t tile.
__global__ __launch_bounds__(256) void k_compress(
    const double* __restrict__ Rbar, int64_t nI, int64_t nJ, int kL, int kR, int dim, int64_t nT,
    double* __restrict__ Gc, int64_t ldc) {
    __shared__ double tile[32][33];
    const int64_t jt = (nJ + 31) / 32;
    const int64_t I = blockIdx.x / jt, J0 = (blockIdx.x - I * jt) * 32;
    const int64_t l = blockIdx.y, t0 = (int64_t)blockIdx.z * 32;
    int64_t nIL = 1, nIR = 1;
    for (int j = 0; j < kL; j++) nIL *= dim;
    for (int j = 0; j < kR; j++) nIR *= dim;
    int lo[3] = {0, 0, 0}, ro[3] = {0, 0, 0};
    tuple_orbs(l, dim, kL, lo);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t t = t0 + tx;
    if (t < nT) tuple_orbs(t, dim, kR, ro);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int jj = ty + 8 * q;
        const int64_t J = J0 + jj;
        if (t < nT && J < nJ)
            tile[jj][tx] = compress_sum(Rbar + (I * nJ + J) * nIL * nIR, lo, ro, kL, kR, dim, nIR);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int tt = ty + 8 * q;
        const int64_t J = J0 + tx;
        if (t0 + tt < nT && J < nJ) Gc[(l * nI + I) * ldc + (t0 + tt) * nJ + J] = tile[tx][tt];
    }
}

// Transpose of k_gather through the side table: for each source determinant
// S (state space, mask Sm[S]) and state s,
//   zbar[S, s] = sum_t sign(K, t) * Ybar[K(S,t), t, s]
// where t runs over the tuples the side string removes from S (annihilators)
// or adds to it (creators), and K is their image in the K space.  Each
// output is owned by one thread: deterministic, no atomics.
// One (tuple of K ascending orbitals o[0] < o[1] < o[2]) term of the side
// adjoint: sign of applying the string (slot order) to src, K's rank, t's rank.
// Encoded as sign * (row + 1) with row the Ybar row (K, t) in the given
// layout, 0 where K falls outside the K space.
template <int K>
__device__ __forceinline__ int64_t side_code(const SidePlan& sd, uint64_t src, const int* o, int n_orbs,
                                            uint64_t kcore, int kspace_core, int nK_val, int dim,
                                            int64_t nK, int64_t nT, int layout) {
    uint64_t tm = 0;
#pragma unroll
    for (int j = 0; j < K; j++) tm |= 1ull << o[j];
    const uint64_t km = sd.is_c ? (src | tm) : (src & ~tm);
    if ((km & kcore) != kcore) return 0;
    uint64_t cur = src;
    int sg = 1;
#pragma unroll
    for (int j = K - 1; j >= 0; j--) {
        const int sl = sd.slot[j];
        const int oo = (sl == 0) ? o[0] : ((sl == 1) ? o[1 < K ? 1 : 0] : o[2 < K ? 2 : 0]);
        sg *= d_sign_above(cur, oo);
        cur ^= 1ull << oo;
    }
    const int64_t Kr = d_lex_rank(d_to_valence(km, n_orbs, kspace_core), nK_val);
    const int64_t t = d_lex_rank(tm, dim);
    const int64_t row = layout == 0 ? Kr * nT + t : t * nK + Kr;
    return sg * (row + 1);
}

template <int K>
__device__ __forceinline__ double side_term(const SidePlan& sd, uint64_t src, const int* o, int n_orbs,
                                            uint64_t kcore, int kspace_core, int nK_val, int dim,
                                            int64_t nK, int64_t nT, const double* __restrict__ Ybar,
                                            int64_t nS, int64_t s, int layout) {
    const int64_t e = side_code<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, layout);
    if (!e) return 0.0;
    const double y = Ybar[((e > 0 ? e : -e) - 1) * nS + s];
    return e > 0 ? y : -y;
}

// The codes of side_code for every source determinant S and tuple j (in the
// enumeration order of k_side_adjoint): tab[S * nTup + j].  One thread per S.
template <int K>
__global__ void k_side_index(SidePlan sd, const uint64_t* __restrict__ Sm, int64_t nS_cfg, int n_orbs,
                             int kspace_core, int64_t nK, int64_t nT, int layout, int64_t nTup,
                             int32_t* __restrict__ tab) {
    XR_GRID_STRIDE(S, nS_cfg) {
        const int dim = 2 * n_orbs;
        const uint64_t src = Sm[S];
        const uint64_t full = (dim >= 64) ? ~0ull : ((1ull << dim) - 1);
        const uint64_t pool = sd.is_c ? (~src & full) : src;
        const uint64_t kcore = d_core_mask(n_orbs, kspace_core);
        const int nK_val = dim - 2 * kspace_core;
        int32_t* out = tab + S * nTup;
        int64_t j = 0;
        int o[3] = {0, 0, 0};
        for (uint64_t p1 = pool; p1; p1 &= p1 - 1) {
            o[0] = __ffsll((unsigned long long)p1) - 1;
            if (K == 1) {
                if (j < nTup) out[j++] = (int32_t)side_code<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, layout);
                continue;
            }
            for (uint64_t p2 = p1 & (p1 - 1); p2; p2 &= p2 - 1) {
                o[1] = __ffsll((unsigned long long)p2) - 1;
                if (K == 2) {
                    if (j < nTup) out[j++] = (int32_t)side_code<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, layout);
                    continue;
                }
                for (uint64_t p3 = p2 & (p2 - 1); p3; p3 &= p3 - 1) {
                    o[2] = __ffsll((unsigned long long)p3) - 1;
                    if (j < nTup) out[j++] = (int32_t)side_code<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, layout);
                }
            }
        }
        for (; j < nTup; j++) out[j] = 0;
    }
}

// zbarT[S, s] = sum_j sign_j Ybar[row_j, s] over the tuples of S, in the
// tuple order of k_side_adjoint (same sums, same order); the codes of one S
// are shared by the lanes of its states.
__global__ __launch_bounds__(256) void k_side_adjoint_tab(const int32_t* __restrict__ tab, int64_t nS_cfg,
                                                          int64_t nTup, const double* __restrict__ Ybar,
                                                          int64_t nS, double* __restrict__ zbarT) {
    XR_GRID_STRIDE(i, nS_cfg * nS) {
        const int64_t S = i / nS, s = i - S * nS;
        const int32_t* e = tab + S * nTup;
        double acc = 0.0;
#pragma unroll 8
        for (int64_t j = 0; j < nTup; j++) {
            const int32_t v = e[j];
            if (v) {
                const double y = Ybar[(int64_t)((v > 0 ? v : -v) - 1) * nS + s];
                acc += v > 0 ? y : -y;
            }
        }
        zbarT[i] = acc;
    }
}

// Each output (S, s) is owned by one thread (deterministic, no atomics); the
// state index s is fastest, so the threads of a wave share S and walk the
// same tuples in lockstep while reading consecutive s.  Tuples are
// enumerated as nested ascending bit loops over the orbitals the string may
// touch, in the lexicographic order of (o[0], o[1], o[2]).
template <int K>
__global__ void k_side_adjoint(SidePlan sd, const uint64_t* __restrict__ Sm, int64_t nS_cfg,
                               int n_orbs, int n_core, int kspace_core, int64_t nK, int64_t nT,
                               const double* __restrict__ Ybar, int64_t nS, int layout,
                               double* __restrict__ zbarT) {
    XR_GRID_STRIDE(i, nS_cfg * nS) {
        const int dim = 2 * n_orbs;
        const int64_t S = i / nS, s = i - (i / nS) * nS;
        const uint64_t src = Sm[S];
        const uint64_t full = (dim >= 64) ? ~0ull : ((1ull << dim) - 1);
        const uint64_t pool = sd.is_c ? (~src & full) : src;
        const uint64_t kcore = d_core_mask(n_orbs, kspace_core);
        const int nK_val = dim - 2 * kspace_core;
        double acc = 0.0;
        int o[3] = {0, 0, 0};
        for (uint64_t p1 = pool; p1; p1 &= p1 - 1) {
            o[0] = __ffsll((unsigned long long)p1) - 1;
            if (K == 1) {
                acc += side_term<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, Ybar, nS, s, layout);
                continue;
            }
            for (uint64_t p2 = p1 & (p1 - 1); p2; p2 &= p2 - 1) {
                o[1] = __ffsll((unsigned long long)p2) - 1;
                if (K == 2) {
                    acc += side_term<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, Ybar, nS, s, layout);
                    continue;
                }
                for (uint64_t p3 = p2 & (p2 - 1); p3; p3 &= p3 - 1) {
                    o[2] = __ffsll((unsigned long long)p3) - 1;
                    acc += side_term<K>(sd, src, o, n_orbs, kcore, kspace_core, nK_val, dim, nK, nT, Ybar, nS, s, layout);
                }
            }
        }
        zbarT[S * nS + s] = acc;
    }
}

// zbar[s * nC + c] += zbarT[c * nS + s]
__global__ void k_add_transpose(const double* __restrict__ zT, int64_t nS, int64_t nC,
                                double* __restrict__ z) {
    __shared__ double tile[32][33];
    int64_t c0 = (int64_t)blockIdx.x * 32, s0 = (int64_t)blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;
    for (int j = ty; j < 32; j += 8) {
        int64_t c = c0 + j, s = s0 + tx;
        if (s < nS && c < nC) tile[j][tx] = zT[c * nS + s];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        int64_t s = s0 + j, c = c0 + tx;
        if (s < nS && c < nC) z[s * nC + c] += tile[tx][j];
    }
}

// S = G + G^T for a square n x n matrix (32 x 32 tiles, the transposed tile
// through LDS).
__global__ void k_sym_sum(const double* __restrict__ G, int64_t n, double* __restrict__ S) {
    __shared__ double tile[32][33];
    const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int j = ty; j < 32; j += 8) {            // tile[j][tx] = G[c0 + j][r0 + tx]
        const int64_t r = c0 + j, col = r0 + tx;
        if (r < n && col < n) tile[j][tx] = G[r * n + col];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {            // S[r0 + j][c0 + tx] = G[r0 + j][c0 + tx] + G[c0 + tx][r0 + j]
        const int64_t r = r0 + j, col = c0 + tx;
        if (r < n && col < n) S[r * n + col] = G[r * n + col] + tile[tx][j];
    }
}

static int side_adjoint(xr_ctx* c, const SidePlan& sd, int n_orbs, int n_core, int kspace_core,
                        int n_elec_src, int64_t nK, int64_t nT, const double* Ybar, int64_t nS,
                        int layout, double* zbar, double* tmpT) {
    const uint64_t* Sm = nullptr;
    int64_t nS_cfg = 0;
    int rc;
    if (sd.k < 1 || sd.k > 3) return xr_fail(c, XR_EINVAL, "side string of 1 to 3 operators expected");
    if ((rc = xr_space_masks(c, n_orbs, n_core, n_elec_src, &Sm, &nS_cfg))) return rc;
    int64_t n = nS_cfg * nS;
    // index table of the (S, tuple) terms: built once per side string and
    // spaces, then every call is a signed gather-sum
    const int dim = 2 * n_orbs;
    const int64_t pool = sd.is_c ? dim - n_elec_src : n_elec_src;
    const int64_t nTup = h_binom((int)pool, sd.k);
    const std::array<int64_t, 12> key{sd.k, sd.is_c, sd.slot[0], sd.slot[1], sd.slot[2], n_orbs, n_core,
                                      kspace_core, n_elec_src, layout, nK, nT};
    int32_t* tab = nullptr;
    auto it = c->sidetab.find(key);
    if (it != c->sidetab.end()) {
        tab = it->second.first;
    } else if (nK * nT < 0x7fffffffLL && nS_cfg * nTup > 0 && nS_cfg * nTup <= (1LL << 28)) {
        XR_HIP(c, hipMalloc(&tab, nS_cfg * nTup * sizeof(int32_t)));
        c->sidetab[key] = std::make_pair(tab, nS_cfg * nTup);
        ProfScope prof(c, XR_K_TABLE, 0.0, 4.0 * nS_cfg * nTup);
        if (sd.k == 1)
            k_side_index<1><<<xr_grid(nS_cfg), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, kspace_core, nK, nT, layout, nTup, tab);
        else if (sd.k == 2)
            k_side_index<2><<<xr_grid(nS_cfg), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, kspace_core, nK, nT, layout, nTup, tab);
        else
            k_side_index<3><<<xr_grid(nS_cfg), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, kspace_core, nK, nT, layout, nTup, tab);
        XR_LAUNCH_CHECK(c);
    }
    if (tab) {
        ProfScope prof(c, XR_K_GATHER, 0.0, 8.0 * n + 4.0 * nS_cfg * nTup);
        k_side_adjoint_tab<<<xr_grid(n), 256, 0, c->stream>>>(tab, nS_cfg, nTup, Ybar, nS, tmpT);
        XR_LAUNCH_CHECK(c);
    } else {
        ProfScope prof(c, XR_K_GATHER, 0.0, 8.0 * n);
        if (sd.k == 1)
            k_side_adjoint<1><<<xr_grid(n), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, n_core, kspace_core,
                                                                 nK, nT, Ybar, nS, layout, tmpT);
        else if (sd.k == 2)
            k_side_adjoint<2><<<xr_grid(n), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, n_core, kspace_core,
                                                                 nK, nT, Ybar, nS, layout, tmpT);
        else
            k_side_adjoint<3><<<xr_grid(n), 256, 0, c->stream>>>(sd, Sm, nS_cfg, n_orbs, n_core, kspace_core,
                                                                 nK, nT, Ybar, nS, layout, tmpT);
        XR_LAUNCH_CHECK(c);
    }
    dim3 grid((unsigned)ceil_div(nS_cfg, 32), (unsigned)ceil_div(nS, 32));
    k_add_transpose<<<grid, dim3(32, 8), 0, c->stream>>>(tmpT, nS, nS_cfg, zbar);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Vbar[p,q,b,a] += Veb[u,w], Vbar[q,p,b,a] -= .., Vbar[p,q,a,b] -= .., Vbar[q,p,a,b] += ..
// (transpose of k_veff); one thread per (u, w), distinct targets.
__device__ __forceinline__ void k_veff_adjoint_elem(int64_t i, const double* __restrict__ Veb, int dim, int64_t nU,
                               double* __restrict__ Vbar) {
    int64_t u = i / nU, w = i - u * nU;
    uint64_t mu = d_lex_unrank(u, dim, 2), mw = d_lex_unrank(w, dim, 2);
    int64_t p = __ffsll((unsigned long long)mu) - 1;
    int64_t q = 63 - __clzll((long long)mu);
    int64_t a = __ffsll((unsigned long long)mw) - 1;
    int64_t b = 63 - __clzll((long long)mw);
    int64_t D = dim;
    double g = Veb[i];
    auto at = [&](int64_t x0, int64_t x1, int64_t x2, int64_t x3) -> double& {
        return Vbar[((x0 * D + x1) * D + x2) * D + x3];
    };
    at(p, q, b, a) += g;
    at(q, p, b, a) -= g;
    at(p, q, a, b) -= g;
    at(q, p, a, b) += g;
}

__global__ void k_veff_adjoint(const double* __restrict__ Veb, int dim, int64_t nU,
                               double* __restrict__ Vbar) {
    XR_GRID_STRIDE(i, nU * nU) k_veff_adjoint_elem(i, Veb, dim, nU, Vbar);
}

// Wgiven: W = Veff Y as xr_monomer_2e_keep left it (same V and ket), used
// instead of recomputing the product; nullptr: recomputed.
static int monomer_2e_adjoint_impl(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI,
                                   int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                                   const double* zket, const double* V, const double* Ebar,
                                   double* zbar_bra, double* zbar_ket, double* Vbar,
                                   const double* Wgiven, int64_t w_len) {
    OpPlan pl;
    xr_plan_op("ccaa", &pl);
    int nb = 0;
    int rc = check_spaces(c, pl, n_orbs, n_core, n_elec, nP, nQ, &nb);
    if (rc) return rc;
    if (nI <= 0 || nJ <= 0 || n_elec < 2) return XR_OK;
    const int dim = 2 * n_orbs;
    double* zTb = (double*)xr_workspace(c, WS_ZTB, nP * nI * sizeof(double));
    double* zTk = (double*)xr_workspace(c, WS_ZTK, nQ * nJ * sizeof(double));
    if (!zTb || !zTk) return xr_fail(c, XR_ENOMEM, "workspace (coefficients)");
    if ((rc = xr_transpose(c, zbra, nI, nP, zTb))) return rc;
    if ((rc = xr_transpose(c, zket, nJ, nQ, zTk))) return rc;
    const uint64_t* Km = nullptr;
    int64_t nK = 0;
    if ((rc = xr_space_masks(c, n_orbs, 0, n_elec - 2, &Km, &nK))) return rc;
    const int64_t nU = h_binom(dim, 2);
    int32_t* tabL = nullptr;
    int32_t* tabR = nullptr;
    double* X = (double*)xr_workspace(c, WS_X, nU * nK * nI * sizeof(double));
    double* Y = (double*)xr_workspace(c, WS_Y, nU * nK * nJ * sizeof(double));
    if (Wgiven && w_len != nU * nK * nJ) return xr_fail(c, XR_ESHAPE, "sigma factor buffer size");
    double* W = Wgiven ? const_cast<double*>(Wgiven) : (double*)xr_workspace(c, WS_W, nU * nK * nJ * sizeof(double));
    double* Ve = (double*)xr_workspace(c, WS_VE, nU * nU * sizeof(double));
    double* Xbar = (double*)xr_workspace(c, WS_XB, nU * nK * nI * sizeof(double));
    double* Ybar = (double*)xr_workspace(c, WS_YB, nU * nK * nJ * sizeof(double));
    double* Wbar = (double*)xr_workspace(c, WS_WB, nU * nK * nJ * sizeof(double));
    double* Veb = (double*)xr_workspace(c, WS_VEB, nU * nU * sizeof(double));
    int64_t nmax = nP * nI > nQ * nJ ? nP * nI : nQ * nJ;
    double* zbT = (double*)xr_workspace(c, WS_ZBT, nmax * sizeof(double));
    if (!X || !Y || !W || !Ve || !Xbar || !Ybar || !Wbar || !Veb || !zbT)
        return xr_fail(c, XR_ENOMEM, "workspace (sigma adjoint)");
    if ((rc = xr_side_table(c, pl.L, Km, nK, n_orbs, n_core, &tabL))) return rc;
    if ((rc = xr_side_table(c, pl.R, Km, nK, n_orbs, n_core, &tabR))) return rc;
    // A block with itself into one adjoint buffer: both sides remove two
    // electrons from the same states, and their tables differ only by the
    // order of the two annihilators (sg = -1 when the slots are swapped), so
    // X = sg Y.  When the caller guarantees a Hermitian two-body part
    // (V[pqrs] = V[rspq], xr_assume_hermitian_v), Ve is symmetric, hence
    // Ybar = Ve^T Wbar = sg W Ebar with W = Ve Y, side_R = sg side_L, and the
    // two side adjoints add up to side_L(W (Ebar + Ebar^T)): one gather, no
    // Ve^T Wbar product, one side adjoint.
    const double sg = (pl.L.slot[0] == pl.R.slot[0]) ? 1.0 : -1.0;
    const bool mirror = c->hermitian_v && zbra == zket && nI == nJ && nP == nQ && zbar_bra == zbar_ket && zbar_ket &&
                        pl.L.k == 2 && pl.R.k == 2 && pl.L.is_c == pl.R.is_c &&
                        pl.L.slot[0] != pl.L.slot[1] && pl.R.slot[0] != pl.R.slot[1];
    if (!mirror && (rc = xr_gather(c, tabL, nK, nU, zTb, nI, X, 1))) return rc;   // X[u][K][I]
    if ((rc = xr_gather(c, tabR, nK, nU, zTk, nJ, Y, 1))) return rc;   // Y[w][K][J]
    k_veff<<<xr_grid(nU * nU), 256, 0, c->stream>>>(V, dim, nU, Ve);
    XR_LAUNCH_CHECK(c);
    const int64_t KJ = nK * nJ;
    // W = Ve Y (nU x KJ)
    if (!Wgiven && (rc = xr_gemm_2e(c, 0, 0, nU, KJ, nU, 1.0, Ve, nU, 0, Y, KJ, 0, 0.0, W, KJ, 0, 1))) return rc;
    if (mirror) {
        double* Es = (double*)xr_workspace(c, WS_GS, nI * nI * sizeof(double));
        if (!Es) return xr_fail(c, XR_ENOMEM, "workspace (sigma adjoint)");
        dim3 grid((unsigned)ceil_div(nI, 32), (unsigned)ceil_div(nI, 32));
        k_sym_sum<<<grid, dim3(32, 8), 0, c->stream>>>(Ebar, nI, Es);   // Es = Ebar + Ebar^T
        XR_LAUNCH_CHECK(c);
        // Xbar[(u,K), I] = sum_J W[(u,K), J] Es[I, J]
        if ((rc = xr_gemm_2e(c, 0, 1, nU * nK, nI, nJ, 1.0, W, nJ, 0, Es, nJ, 0, 0.0, Xbar, nI, 0, 1)))
            return rc;
        if (Vbar) {
            // Wbar = X Ebar = sg Y Ebar, Veb = Wbar Y^T, then scatter into Vbar
            if ((rc = xr_gemm_2e(c, 0, 0, nU * nK, nJ, nI, sg, Y, nI, 0, Ebar, nJ, 0, 0.0, Wbar, nJ, 0, 1)))
                return rc;
            if ((rc = xr_gemm_2e(c, 0, 1, nU, nU, KJ, 1.0, Wbar, KJ, 0, Y, KJ, 0, 0.0, Veb, nU, 0, 1)))
                return rc;
            k_veff_adjoint<<<xr_grid(nU * nU), 256, 0, c->stream>>>(Veb, dim, nU, Vbar);
            XR_LAUNCH_CHECK(c);
        }
        return side_adjoint(c, pl.L, n_orbs, n_core, 0, n_elec, nK, nU, Xbar, nI, 1, zbar_bra, zbT);
    }
    // Xbar[(u,K), I] = sum_J W[(u,K), J] Ebar[I, J]    ((nU nK) x nI)
    if ((rc = xr_gemm_2e(c, 0, 1, nU * nK, nI, nJ, 1.0, W, nJ, 0, Ebar, nJ, 0, 0.0, Xbar, nI, 0, 1)))
        return rc;
    // Wbar[(u,K), J] = sum_I X[(u,K), I] Ebar[I, J]
    if ((rc = xr_gemm_2e(c, 0, 0, nU * nK, nJ, nI, 1.0, X, nI, 0, Ebar, nJ, 0, 0.0, Wbar, nJ, 0, 1)))
        return rc;
    // Ybar = Ve^T Wbar (nU x KJ)
    if ((rc = xr_gemm_2e(c, 1, 0, nU, KJ, nU, 1.0, Ve, nU, 0, Wbar, KJ, 0, 0.0, Ybar, KJ, 0, 1)))
        return rc;
    if (Vbar) {
        // Veb = Wbar Y^T (nU x nU), then scatter into Vbar
        if ((rc = xr_gemm_2e(c, 0, 1, nU, nU, KJ, 1.0, Wbar, KJ, 0, Y, KJ, 0, 0.0, Veb, nU, 0, 1)))
            return rc;
        k_veff_adjoint<<<xr_grid(nU * nU), 256, 0, c->stream>>>(Veb, dim, nU, Vbar);
        XR_LAUNCH_CHECK(c);
    }
    if (zbar_ket)
        if ((rc = side_adjoint(c, pl.R, n_orbs, n_core, 0, n_elec, nK, nU, Ybar, nJ, 1, zbar_ket, zbT)))
            return rc;
    if (zbar_bra)
        if ((rc = side_adjoint(c, pl.L, n_orbs, n_core, 0, n_elec, nK, nU, Xbar, nI, 1, zbar_bra, zbT)))
            return rc;
    return XR_OK;
}

extern "C" int xr_monomer_2e_adjoint(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI,
                                     int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                                     const double* zket, const double* V, const double* Ebar,
                                     double* zbar_bra, double* zbar_ket, double* Vbar) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    return monomer_2e_adjoint_impl(c, n_orbs, n_core, n_elec, nI, nP, zbra, nJ, nQ, zket, V, Ebar, zbar_bra,
                                   zbar_ket, Vbar, nullptr, 0);
}

extern "C" int xr_monomer_2e_adjoint_keep(xr_ctx* c, int n_orbs, int n_core, int n_elec, int64_t nI,
                                          int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                                          const double* zket, const double* V, const double* Ebar,
                                          double* zbar_bra, double* zbar_ket, double* Vbar,
                                          const double* W, int64_t w_len) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!W) return xr_fail(c, XR_EINVAL, "null sigma factor buffer");
    return monomer_2e_adjoint_impl(c, n_orbs, n_core, n_elec, nI, nP, zbra, nJ, nQ, zket, V, Ebar, zbar_bra,
                                   zbar_ket, Vbar, W, w_len);
}

extern "C" int xr_density_adjoint(xr_ctx* c, const char* ops, int n_orbs, int n_core,
                                  int n_elec_ket, int64_t nI, int64_t nP, const double* zbra,
                                  int64_t nJ, int64_t nQ, const double* zket, const double* Rbar,
                                  double* zbar_bra, double* zbar_ket) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    OpPlan pl;
    if (!xr_plan_op(ops, &pl)) return xr_fail(c, XR_EUNSUPPORTED, "operator string not supported");
    int nb = 0;
    int rc = check_spaces(c, pl, n_orbs, n_core, n_elec_ket, nP, nQ, &nb);
    if (rc) return rc;
    if (nI <= 0 || nJ <= 0) return XR_OK;
    if (pl.L.k && pl.R.k && n_elec_ket - pl.R.k < 0) return XR_OK;  // density identically zero
    const int dim = 2 * n_orbs;
    Factors f;
    if ((rc = prepare_factors(c, pl, n_orbs, n_core, n_elec_ket, nb, nI, nP, zbra, nJ, nQ, zket, &f)))
        return rc;
    const int64_t Mg = f.nLt * nI, Ng = f.nT * nJ;
    double* Gc = (double*)xr_workspace(c, WS_GC, Mg * Ng * sizeof(double));
    double* Xbar = (double*)xr_workspace(c, WS_XB, f.nK * Mg * sizeof(double));
    double* Ybar = (double*)xr_workspace(c, WS_YB, f.nK * Ng * sizeof(double));
    int64_t nmax = nP * nI > nQ * nJ ? nP * nI : nQ * nJ;
    double* zbT = (double*)xr_workspace(c, WS_ZBT, nmax * sizeof(double));
    if (!Gc || !Xbar || !Ybar || !zbT) return xr_fail(c, XR_ENOMEM, "workspace (adjoint)");
    {
        int64_t n = Mg * Ng;
        ProfScope prof(c, XR_K_EXPAND, 0.0, 8.0 * n * (pl.k > 2 ? 4 : 2));
        if (nI * ceil_div(nJ, 32) > 0x7fffffffLL || f.nLt > 65535 || ceil_div(f.nT, 32) > 65535)
            return xr_fail(c, XR_EUNSUPPORTED, "density adjoint grid too large");
        dim3 grid((unsigned)(nI * ceil_div(nJ, 32)), (unsigned)f.nLt, (unsigned)ceil_div(f.nT, 32));
        k_compress<<<grid, 256, 0, c->stream>>>(Rbar, nI, nJ, pl.L.k, pl.R.k, dim, f.nT, Gc, Ng);
        XR_LAUNCH_CHECK(c);
    }
    // One-body string 'ca' between a block and itself, adjoints into one
    // buffer: X = Y (both sides remove one electron from the same states,
    // through the same table) and the two side adjoints are the same map, so
    // side(Y Gc^T) + side(X Gc) = side(X (Gc + Gc^T)): one GEMM, one side
    // adjoint.
    const bool mirror = ops[0] == 'c' && ops[1] == 'a' && ops[2] == 0 && zbra == zket && nI == nJ &&
                        nP == nQ && Mg == Ng && zbar_bra == zbar_ket && zbar_ket && pl.R.k;
    if (mirror) {
        double* Gs = (double*)xr_workspace(c, WS_GS, Mg * Ng * sizeof(double));
        if (!Gs) return xr_fail(c, XR_ENOMEM, "workspace (adjoint)");
        dim3 grid((unsigned)ceil_div(Mg, 32), (unsigned)ceil_div(Mg, 32));
        k_sym_sum<<<grid, dim3(32, 8), 0, c->stream>>>(Gc, Mg, Gs);   // Gs = Gc + Gc^T
        XR_LAUNCH_CHECK(c);
        if ((rc = xr_dgemm_launch(c, 0, 0, f.nK, Ng, Mg, 1.0, f.X, Mg, 0, Gs, Ng, 0, 0.0, Ybar, Ng, 0, 1)))
            return rc;
        return side_adjoint(c, pl.R, n_orbs, n_core, f.kspace_core, n_elec_ket, f.nK, f.nT, Ybar, nJ, 0,
                            zbar_ket, zbT);
    }
    // Ybar = X Gc  (nK x Ng);  Xbar = Y Gc^T  (nK x Mg)
    if ((rc = xr_dgemm_launch(c, 0, 0, f.nK, Ng, Mg, 1.0, f.X, Mg, 0, Gc, Ng, 0, 0.0, Ybar, Ng, 0, 1)))
        return rc;
    if ((rc = xr_dgemm_launch(c, 0, 1, f.nK, Mg, Ng, 1.0, f.Y, Ng, 0, Gc, Ng, 0, 0.0, Xbar, Mg, 0, 1)))
        return rc;
    if (zbar_ket) {
        if (pl.R.k) {
            if ((rc = side_adjoint(c, pl.R, n_orbs, n_core, f.kspace_core, n_elec_ket, f.nK, f.nT, Ybar,
                                   nJ, 0, zbar_ket, zbT)))
                return rc;
        } else {
            // Y = zket^T, so zbar_ket[J, Q] += Ybar[Q, J]
            dim3 grid((unsigned)ceil_div(nQ, 32), (unsigned)ceil_div(nJ, 32));
            k_add_transpose<<<grid, dim3(32, 8), 0, c->stream>>>(Ybar, nJ, nQ, zbar_ket);
            XR_LAUNCH_CHECK(c);
        }
    }
    if (zbar_bra) {
        if (pl.L.k) {
            if ((rc = side_adjoint(c, pl.L, n_orbs, n_core, f.kspace_core, nb, f.nK, f.nLt, Xbar, nI, 0,
                                   zbar_bra, zbT)))
                return rc;
        } else {
            dim3 grid((unsigned)ceil_div(nP, 32), (unsigned)ceil_div(nI, 32));
            k_add_transpose<<<grid, dim3(32, 8), 0, c->stream>>>(Xbar, nI, nP, zbar_bra);
            XR_LAUNCH_CHECK(c);
        }
    }
    return XR_OK;
}
