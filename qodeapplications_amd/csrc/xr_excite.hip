// Excitation lists in the reference loop order (integer work, bit-exact).
//
// For a ket determinant Q (64-bit occupation mask) the operator string is
// applied right to left.  At each level the candidate spin orbitals are the
// occupied ones for an annihilator and the empty ones for a creator, visited
// in ascending order, which reproduces the nested index loops of
// general-XRCC/density_tensors.c:162-556 (loops over all dim indices with the
// Pauli-violating ones skipped).  The sign of each step is the parity of the
// occupied orbitals above the index.
#include "xr_common.h"
#include <hipcub/hipcub.hpp>

struct OpString {
    int k;
    int is_c[XR_MAX_OPS];  // 1 = creation
    int delta;             // change of electron number
};

static int parse_ops(const char* ops, OpString* s) {
    if (!ops) return 0;
    int k = 0;
    int d = 0;
    for (const char* p = ops; *p; p++) {
        if (k >= XR_MAX_OPS) return 0;
        if (*p == 'c') { s->is_c[k++] = 1; d++; }
        else if (*p == 'a') { s->is_c[k++] = 0; d--; }
        else return 0;
    }
    if (k == 0) return 0;
    s->k = k;
    s->delta = d;
    return 1;
}

template <bool FILL>
__global__ void k_excite(OpString os, int dim, int n_orbs, int n_core, const uint64_t* ket,
                         int64_t q0, int64_t q1, int64_t* counts, const int64_t* offsets,
                         int64_t* e_q, int64_t* e_idx, int64_t* e_p, int8_t* e_sign) {
  for (int64_t Q = q0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; Q < q1;
       Q += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t full = (dim >= 64) ? ~0ull : ((1ull << dim) - 1);
    const uint64_t core = d_core_mask(n_orbs, n_core);
    const int nval = dim - 2 * n_core;
    uint64_t occ_in[XR_MAX_OPS], cand[XR_MAX_OPS];
    int sg_in[XR_MAX_OPS], ind[XR_MAX_OPS];
    int64_t pos = FILL ? offsets[Q - q0] : 0;
    int64_t cnt = 0;
    int L = os.k - 1;
    occ_in[L] = ket[Q];
    sg_in[L] = 1;
    cand[L] = os.is_c[L] ? (~occ_in[L] & full) : occ_in[L];
    while (true) {
        if (cand[L] == 0) {
            L++;
            if (L >= os.k) break;
            continue;
        }
        int x = __ffsll((unsigned long long)cand[L]) - 1;
        cand[L] &= cand[L] - 1;
        int s = sg_in[L] * d_sign_above(occ_in[L], x);
        uint64_t nxt = occ_in[L] ^ (1ull << x);
        ind[L] = x;
        if (L > 0) {
            occ_in[L - 1] = nxt;
            sg_in[L - 1] = s;
            cand[L - 1] = os.is_c[L - 1] ? (~nxt & full) : nxt;
            L--;
            continue;
        }
        if ((nxt & core) != core) continue;
        if (FILL) {
            int64_t idx = 0;
            for (int l = 0; l < os.k; l++) idx = idx * dim + ind[l];
            e_q[pos] = Q;
            e_idx[pos] = idx;
            e_p[pos] = d_lex_rank(d_to_valence(nxt, n_orbs, n_core), nval);
            e_sign[pos] = (int8_t)s;
            pos++;
        } else {
            cnt++;
        }
    }
    if (!FILL) counts[Q - q0] = cnt;
  }
}

static int check_common(xr_ctx* c, const char* ops, int n_orbs, int n_core, int n_elec_ket,
                        OpString* os) {
    if (!parse_ops(ops, os)) return xr_fail(c, XR_EUNSUPPORTED, "operator string must be 1-5 of 'c'/'a'");
    if (n_orbs <= 0 || 2 * n_orbs > XR_MAX_SO) return xr_fail(c, XR_EUNSUPPORTED, "need 0 < 2*n_orbs <= 64");
    if (n_core < 0 || n_core > n_orbs) return xr_fail(c, XR_EINVAL, "bad n_core");
    int nb = n_elec_ket + os->delta;
    if (n_elec_ket < 2 * n_core || n_elec_ket > 2 * n_orbs || nb < 2 * n_core || nb > 2 * n_orbs)
        return xr_fail(c, XR_EINVAL, "electron counts incompatible with orbitals / core");
    return XR_OK;
}

extern "C" int xr_excitation_count(xr_ctx* c, const char* ops, int n_orbs, int n_core,
                                   int n_elec_ket, const uint64_t* d_ket, int64_t q0, int64_t q1,
                                   int64_t* d_offsets, int64_t* total) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    OpString os;
    int rc = check_common(c, ops, n_orbs, n_core, n_elec_ket, &os);
    if (rc) return rc;
    if (q0 < 0 || q1 < q0) return xr_fail(c, XR_EINVAL, "bad ket range");
    int64_t n = q1 - q0;
    int64_t* counts = (int64_t*)xr_workspace(c, 14, (n + 1) * sizeof(int64_t));
    if (!counts) return xr_fail(c, XR_ENOMEM, "workspace");
    XR_HIP(c, hipMemsetAsync(counts, 0, (n + 1) * sizeof(int64_t), c->stream));
    if (n > 0) {
        k_excite<false><<<xr_grid(n, 128), 128, 0, c->stream>>>(
            os, 2 * n_orbs, n_orbs, n_core, d_ket, q0, q1, counts, nullptr, nullptr, nullptr,
            nullptr, nullptr);
        XR_LAUNCH_CHECK(c);
    }
    size_t tmp_bytes = 0;
    XR_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, d_offsets, n + 1,
                                               c->stream));
    void* tmp = xr_workspace(c, 13, tmp_bytes);
    if (!tmp) return xr_fail(c, XR_ENOMEM, "workspace");
    XR_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, d_offsets, n + 1,
                                               c->stream));
    int64_t t = 0;
    XR_HIP(c, hipMemcpyAsync(&t, d_offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    if (total) *total = t;
    return XR_OK;
}

extern "C" int xr_excitation_fill(xr_ctx* c, const char* ops, int n_orbs, int n_core,
                                  int n_elec_ket, const uint64_t* d_ket, int64_t q0, int64_t q1,
                                  const int64_t* d_offsets, int64_t* d_q, int64_t* d_idx,
                                  int64_t* d_p, int8_t* d_sign) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    OpString os;
    int rc = check_common(c, ops, n_orbs, n_core, n_elec_ket, &os);
    if (rc) return rc;
    if (q0 < 0 || q1 < q0) return xr_fail(c, XR_EINVAL, "bad ket range");
    int64_t n = q1 - q0;
    if (n == 0) return XR_OK;
    k_excite<true><<<xr_grid(n, 128), 128, 0, c->stream>>>(
        os, 2 * n_orbs, n_orbs, n_core, d_ket, q0, q1, nullptr, d_offsets, d_q, d_idx, d_p,
        d_sign);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}
