// H_contractions.c families evaluated for all state pairs at once.
//
// Every reference contraction is a bilinear form  R1[a] . W . R2[b]  in the
// densities of two fragments, with W an integral block whose axes are
// permuted to match the density index order.  Over all state pairs a, b it
// becomes two FP64 GEMMs:  G = R1 * W  (nA x k2), then  out = G * R2^T.
#include "xr_common.h"
#include "xr_gemm.h"
#include <vector>
#include <map>
#include <mutex>

// out[o0,o1,o2,o3] = in[...] with out axis j = in axis perm[j]
__device__ __forceinline__ void k_permute4_elem(int64_t i, const double* __restrict__ in, int64_t d0, int64_t d1, int64_t d2,
                           int64_t d3, int p0, int p1, int p2, int p3, double* __restrict__ out) {
    int64_t din[4] = {d0, d1, d2, d3};
    int p[4] = {p0, p1, p2, p3};
    int64_t dout[4] = {din[p[0]], din[p[1]], din[p[2]], din[p[3]]};
    int64_t o[4];
    int64_t x = i;
    for (int j = 3; j >= 0; j--) {
        o[j] = x % dout[j];
        x /= dout[j];
    }
    int64_t ix[4];
    for (int j = 0; j < 4; j++) ix[p[j]] = o[j];
    out[i] = in[((ix[0] * d1 + ix[1]) * d2 + ix[2]) * d3 + ix[3]];
}

__global__ void k_permute4(const double* __restrict__ in, int64_t d0, int64_t d1, int64_t d2,
                           int64_t d3, int p0, int p1, int p2, int p3, double* __restrict__ out) {
    int64_t n = d0 * d1 * d2 * d3;
    XR_GRID_STRIDE(i, n) k_permute4_elem(i, in, d0, d1, d2, d3, p0, p1, p2, p3, out);
}

static int permute4(xr_ctx* c, const double* in, int64_t d0, int64_t d1, int64_t d2, int64_t d3,
                    int p0, int p1, int p2, int p3, double* out) {
    int64_t n = d0 * d1 * d2 * d3;
    if (n <= 0) return XR_OK;
    k_permute4<<<xr_grid(n), 256, 0, c->stream>>>(in, d0, d1, d2, d3, p0, p1, p2, p3, out);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

enum { WS_CW1 = 20, WS_CW2, WS_CG1, WS_CG2 };

#define XR_ENTRY(c)            \
    if (!(c)) return XR_EINVAL; \
    std::lock_guard<std::mutex> _g((c)->mu)

extern "C" int xr_monomer_1e(xr_ctx* c, int64_t n, int64_t nA, const double* Rca, const double* h,
                             double* out) {
    XR_ENTRY(c);
    if (n <= 0 || nA < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // out[a] = sum_{pq} Rca[a, pq] h[pq]
    return xr_dgemm_launch(c, 0, 0, nA, 1, n * n, 1.0, Rca, n * n, 0, h, 1, 0, 0.0, out, 1, 0, 1);
}

extern "C" int xr_dimer_2min2pls(xr_ctx* c, int64_t n1, int64_t n2, int64_t nA, const double* Rcc1,
                                 int64_t nB, const double* Raa2, const double* V, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || nA < 0 || nB < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // W[(p1 q1), (s2 r2)] = V[p1, q1, r2, s2]
    double* W = (double*)xr_workspace(c, WS_CW1, n1 * n1 * n2 * n2 * sizeof(double));
    double* G = (double*)xr_workspace(c, WS_CG1, nA * n2 * n2 * sizeof(double));
    if (!W || !G) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    if ((rc = permute4(c, V, n1, n1, n2, n2, 0, 1, 3, 2, W))) return rc;
    if ((rc = xr_gemm_2e(c, 0, 0, nA, n2 * n2, n1 * n1, 1.0, Rcc1, n1 * n1, 0, W, n2 * n2, 0,
                              0.0, G, n2 * n2, 0, 1)))
        return rc;
    return xr_gemm_2e(c, 0, 1, nA, nB, n2 * n2, 1.0, G, n2 * n2, 0, Raa2, n2 * n2, 0, 0.0, out,
                           nB, 0, 1);
}

extern "C" int xr_dimer_ExEx(xr_ctx* c, int64_t n1, int64_t n2, int64_t nA, const double* Rca1,
                             int64_t nB, const double* Rca2, const double* V, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || nA < 0 || nB < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // W[(p1 r1), (q2 s2)] = V[p1, q2, r1, s2]
    double* W = (double*)xr_workspace(c, WS_CW1, n1 * n1 * n2 * n2 * sizeof(double));
    double* G = (double*)xr_workspace(c, WS_CG1, nA * n2 * n2 * sizeof(double));
    if (!W || !G) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    if ((rc = permute4(c, V, n1, n2, n1, n2, 0, 2, 1, 3, W))) return rc;
    if ((rc = xr_gemm_2e(c, 0, 0, nA, n2 * n2, n1 * n1, 1.0, Rca1, n1 * n1, 0, W, n2 * n2, 0,
                              0.0, G, n2 * n2, 0, 1)))
        return rc;
    return xr_gemm_2e(c, 0, 1, nA, nB, n2 * n2, 4.0, G, n2 * n2, 0, Rca2, n2 * n2, 0, 0.0, out,
                           nB, 0, 1);
}

extern "C" int xr_dimer_1min1pls_1e(xr_ctx* c, int64_t n1, int64_t n2, int64_t nA, const double* Rc1,
                                    int64_t nB, const double* Ra2, const double* h, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || nA < 0 || nB < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    double* T = (double*)xr_workspace(c, WS_CG1, nA * n2 * sizeof(double));
    if (!T) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    // T = Rc1 h (nA x n2);  out = T Ra2^T
    if ((rc = xr_dgemm_launch(c, 0, 0, nA, n2, n1, 1.0, Rc1, n1, 0, h, n2, 0, 0.0, T, n2, 0, 1)))
        return rc;
    return xr_dgemm_launch(c, 0, 1, nA, nB, n2, 1.0, T, n2, 0, Ra2, n2, 0, 0.0, out, nB, 0, 1);
}

extern "C" int xr_dimer_1min1pls_2e(xr_ctx* c, int64_t n1, int64_t n2, int64_t nA, const double* Rc1,
                                    const double* Rcca1, int64_t nB, const double* Ra2,
                                    const double* Rcaa2, const double* V1112, const double* V1222,
                                    double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || nA < 0 || nB < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    const int64_t n1c = n1 * n1 * n1, n2c = n2 * n2 * n2;
    double* W1 = (double*)xr_workspace(c, WS_CW1, n1c * n2 * sizeof(double));
    double* W2 = (double*)xr_workspace(c, WS_CW2, n1 * n2c * sizeof(double));
    double* T = (double*)xr_workspace(c, WS_CG1, nA * n2 * sizeof(double));
    double* S = (double*)xr_workspace(c, WS_CG2, nB * n1 * sizeof(double));
    if (!W1 || !W2 || !T || !S) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    // W1[(q1 p1 r1), s2] = V1112[p1, q1, r1, s2];  W2[p1, (q2 s2 r2)] = V1222[p1, q2, r2, s2]
    if ((rc = permute4(c, V1112, n1, n1, n1, n2, 1, 0, 2, 3, W1))) return rc;
    if ((rc = permute4(c, V1222, n1, n2, n2, n2, 0, 1, 3, 2, W2))) return rc;
    // T = Rcca1 W1 (nA x n2);  S[b, p1] = sum Rcaa2[b, (q2 s2 r2)] W2[p1, (q2 s2 r2)]
    if ((rc = xr_gemm_2e(c, 0, 0, nA, n2, n1c, 1.0, Rcca1, n1c, 0, W1, n2, 0, 0.0, T, n2, 0, 1)))
        return rc;
    if ((rc = xr_gemm_2e(c, 0, 1, nB, n1, n2c, 1.0, Rcaa2, n2c, 0, W2, n2c, 0, 0.0, S, n1, 0, 1)))
        return rc;
    // out = 2 (T Ra2^T + Rc1 S^T)
    if ((rc = xr_gemm_2e(c, 0, 1, nA, nB, n2, 2.0, T, n2, 0, Ra2, n2, 0, 0.0, out, nB, 0, 1)))
        return rc;
    return xr_gemm_2e(c, 0, 1, nA, nB, n1, 2.0, Rc1, n1, 0, S, n1, 0, 1.0, out, nB, 0, 1);
}

// dst[r * ldd + j] = src[r * lds + j] for r < nr, j < nc
__global__ void k_copy_cols(const double* __restrict__ src, int64_t lds, int64_t nr, int64_t nc,
                            double* __restrict__ dst, int64_t ldd) {
    XR_GRID_STRIDE(i, nr * nc) {
        const int64_t r = i / nc, j = i - r * nc;
        dst[r * ldd + j] = src[r * lds + j];
    }
}

// The one-electron-transfer block (fragment 1 gains the electron) with its
// one- and two-electron parts in one product:
//   out = [T | Rc1] [Ra2 | S]^T,   T = Rc1 h + 2 Rcca1 W1,   S = 2 Rcaa2 W2^T
// (xr_dimer_1min1pls_1e + xr_dimer_1min1pls_2e, each nA x nB output written
// once instead of three times).
extern "C" int xr_dimer_1min1pls(xr_ctx* c, int64_t n1, int64_t n2, int64_t nA, const double* Rc1,
                                 const double* Rcca1, int64_t nB, const double* Ra2,
                                 const double* Rcaa2, const double* h, const double* V1112,
                                 const double* V1222, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || nA < 0 || nB < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    if (nA == 0 || nB == 0) return XR_OK;
    const int64_t n1c = n1 * n1 * n1, n2c = n2 * n2 * n2, Kc = n1 + n2;
    double* W1 = (double*)xr_workspace(c, WS_CW1, n1c * n2 * sizeof(double));
    double* W2 = (double*)xr_workspace(c, WS_CW2, n1 * n2c * sizeof(double));
    double* Ac = (double*)xr_workspace(c, WS_CG1, nA * Kc * sizeof(double));
    double* Bc = (double*)xr_workspace(c, WS_CG2, nB * Kc * sizeof(double));
    if (!W1 || !W2 || !Ac || !Bc) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    if ((rc = permute4(c, V1112, n1, n1, n1, n2, 1, 0, 2, 3, W1))) return rc;
    if ((rc = permute4(c, V1222, n1, n2, n2, n2, 0, 1, 3, 2, W2))) return rc;
    // Ac[:, :n2] = 2 Rcca1 W1 + Rc1 h;  Ac[:, n2:] = Rc1
    if ((rc = xr_gemm_2e(c, 0, 0, nA, n2, n1c, 2.0, Rcca1, n1c, 0, W1, n2, 0, 0.0, Ac, Kc, 0, 1))) return rc;
    if ((rc = xr_dgemm_launch(c, 0, 0, nA, n2, n1, 1.0, Rc1, n1, 0, h, n2, 0, 1.0, Ac, Kc, 0, 1))) return rc;
    k_copy_cols<<<xr_grid(nA * n1), 256, 0, c->stream>>>(Rc1, n1, nA, n1, Ac + n2, Kc);
    XR_LAUNCH_CHECK(c);
    // Bc[:, :n2] = Ra2;  Bc[:, n2:] = 2 Rcaa2 W2^T
    k_copy_cols<<<xr_grid(nB * n2), 256, 0, c->stream>>>(Ra2, n2, nB, n2, Bc, Kc);
    XR_LAUNCH_CHECK(c);
    if ((rc = xr_gemm_2e(c, 0, 1, nB, n1, n2c, 2.0, Rcaa2, n2c, 0, W2, n2c, 0, 0.0, Bc + n2, Kc, 0, 1))) return rc;
    return xr_gemm_2e(c, 0, 1, nA, nB, Kc, 1.0, Ac, Kc, 0, Bc, Kc, 0, 0.0, out, nB, 0, 1);
}

// out[a, b, c] = f * sum_{x y r s} W[(x y), r, s] R1[a, (x y)] R2[b, r] R3[c, s]
// with W the permuted integral block; the three trimer families differ only
// in the permutation, the prefactor and which densities they take.
static int trimer_chain(xr_ctx* c, const double* V, int64_t d0, int64_t d1, int64_t d2, int64_t d3,
                        int p0, int p1, int p2, int p3, double f, int64_t n1, int64_t n2,
                        int64_t n3, int64_t nA, const double* R1, int64_t nB, const double* R2,
                        int64_t nC, const double* R3, double* out) {
    const int64_t k1 = n1 * n1;
    double* W = (double*)xr_workspace(c, WS_CW1, k1 * n2 * n3 * sizeof(double));
    double* G1 = (double*)xr_workspace(c, WS_CG1, nA * n2 * n3 * sizeof(double));
    double* G2 = (double*)xr_workspace(c, WS_CG2, nA * nB * n3 * sizeof(double));
    if (!W || !G1 || !G2) return xr_fail(c, XR_ENOMEM, "workspace");
    int rc;
    if ((rc = permute4(c, V, d0, d1, d2, d3, p0, p1, p2, p3, W))) return rc;
    // G1[a, (r s)] = sum_{xy} R1[a, xy] W[xy, (r s)]
    if ((rc = xr_gemm_2e(c, 0, 0, nA, n2 * n3, k1, 1.0, R1, k1, 0, W, n2 * n3, 0, 0.0, G1,
                              n2 * n3, 0, 1)))
        return rc;
    // G2[a][b, s] = sum_r R2[b, r] G1[a][r, s]     (batched over a)
    for (int64_t a0 = 0; a0 < nA; a0 += 65535) {
        int64_t nb = (nA - a0 < 65535) ? (nA - a0) : 65535;
        if ((rc = xr_gemm_2e(c, 0, 0, nB, n3, n2, 1.0, R2, n2, 0, G1 + a0 * n2 * n3, n3, n2 * n3,
                                  0.0, G2 + a0 * nB * n3, n3, nB * n3, nb)))
            return rc;
    }
    // out[(a b), c] = f * sum_s G2[(a b), s] R3[c, s]
    return xr_gemm_2e(c, 0, 1, nA * nB, nC, n3, f, G2, n3, 0, R3, n3, 0, 0.0, out, nC, 0, 1);
}

extern "C" int xr_trimer_2min1pls1pls(xr_ctx* c, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                                      const double* Rcc1, int64_t nB, const double* Ra2,
                                      int64_t nC, const double* Ra3, const double* V, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || nA < 0 || nB < 0 || nC < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // W[(q1 p1), r2, s3] = V[p1, q1, r2, s3]   (R1 = Rcc1[a, q1, p1])
    return trimer_chain(c, V, n1, n1, n2, n3, 1, 0, 2, 3, 2.0, n1, n2, n3, nA, Rcc1, nB, Ra2, nC,
                        Ra3, out);
}

extern "C" int xr_trimer_2pls1min1min(xr_ctx* c, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                                      const double* Raa1, int64_t nB, const double* Rc2,
                                      int64_t nC, const double* Rc3, const double* V, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || nA < 0 || nB < 0 || nC < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // W[(q1 p1), r2, s3] = V[r2, s3, p1, q1]   (R1 = Raa1[a, q1, p1])
    return trimer_chain(c, V, n2, n3, n1, n1, 3, 2, 0, 1, 2.0, n1, n2, n3, nA, Raa1, nB, Rc2, nC,
                        Rc3, out);
}

extern "C" int xr_trimer_Ex1min1pls(xr_ctx* c, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                                    const double* Rca1, int64_t nB, const double* Rc2, int64_t nC,
                                    const double* Ra3, const double* V, double* out) {
    XR_ENTRY(c);
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || nA < 0 || nB < 0 || nC < 0) return xr_fail(c, XR_EINVAL, "bad sizes");
    // W[(p1 q1), r2, s3] = V[p1, r2, q1, s3]   (R1 = Rca1[a, p1, q1])
    return trimer_chain(c, V, n1, n2, n1, n3, 0, 2, 1, 3, 4.0, n1, n2, n3, nA, Rca1, nB, Rc2, nC,
                        Ra3, out);
}

// ---------------------------------------------------------------------------
// XR matrix bookkeeping
// ---------------------------------------------------------------------------
struct ScatterDesc {
    int nfrag;
    int64_t nI[3], nJ[3], stride[3], d[3];
    int diag[3];
    const int64_t* rows[3];
    const int64_t* cols[3];
};

__device__ __forceinline__ void k_block_scatter_elem(int64_t i, ScatterDesc sd, const double* __restrict__ blk, double scale,
                                double* __restrict__ out, int64_t ldo, int64_t total) {
    int64_t x = i, src = 0, row = 0, col = 0;
    int64_t I[3], J[3];
    for (int k = sd.nfrag - 1; k >= 0; k--) {
        int64_t nk = sd.diag[k] ? sd.nI[k] : sd.nI[k] * sd.nJ[k];
        int64_t e = x % nk;
        x /= nk;
        if (sd.diag[k]) { I[k] = e; J[k] = e; }
        else { I[k] = e / sd.nJ[k]; J[k] = e % sd.nJ[k]; }
        src += (sd.diag[k] ? 0 : e) * sd.stride[k];
    }
    for (int k = 0; k < sd.nfrag; k++) {
        row = row * sd.d[k] + sd.rows[k][I[k]];
        col = col * sd.d[k] + sd.cols[k][J[k]];
    }
    double v = blk ? scale * blk[src] : scale;
    out[row * ldo + col] += v;
}

__global__ void k_block_scatter(ScatterDesc sd, const double* __restrict__ blk, double scale,
                                double* __restrict__ out, int64_t ldo, int64_t total) {
    XR_GRID_STRIDE(i, total) k_block_scatter_elem(i, sd, blk, scale, out, ldo, total);
}

extern "C" int xr_block_scatter(xr_ctx* c, int nfrag, const int64_t* nI, const int64_t* nJ,
                                const int64_t* stride, const int* diag_only,
                                const int64_t* const* d_rows, const int64_t* const* d_cols,
                                const int64_t* d, const double* d_block, double scale,
                                double* d_out, int64_t ldo) {
    XR_ENTRY(c);
    if (nfrag < 1 || nfrag > 3) return xr_fail(c, XR_EUNSUPPORTED, "1 to 3 fragments per block");
    ScatterDesc sd;
    sd.nfrag = nfrag;
    int64_t total = 1;
    for (int k = 0; k < nfrag; k++) {
        sd.nI[k] = nI[k];
        sd.nJ[k] = nJ[k];
        sd.stride[k] = stride[k];
        sd.d[k] = d[k];
        sd.diag[k] = diag_only[k];
        sd.rows[k] = d_rows[k];
        sd.cols[k] = d_cols[k];
        if (diag_only[k] && nI[k] != nJ[k]) return xr_fail(c, XR_ESHAPE, "diagonal-only fragment needs nI == nJ");
        total *= diag_only[k] ? nI[k] : nI[k] * nJ[k];
    }
    if (total <= 0) return XR_OK;
    ProfScope prof(c, XR_K_SCATTER, 0.0, 16.0 * total);
    k_block_scatter<<<xr_grid(total), 256, 0, c->stream>>>(sd, d_block, scale, d_out, ldo, total);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

#define XR_MAX_FRAG 8
#define XR_MAX_FRAG_AX 8

// ---------------------------------------------------------------------------
// Axis packing for the matrix-free XR operator.  A vector block over the
// tensor-product basis is a row-major tensor src[s_0]...[s_{nf-1}][j]
// (dims[m] states of fragment m, ncol columns).  Packing brings one or two
// fragment axes ("front") to the rows, optionally reordered through rowmap
// (packed row r holds front index rowmap[r] = a * dims[f1] + b), and the
// remaining axes plus the column to the packed columns, in order:
//   packed[r * ld_packed + x],  x = ((i_rest0 * d_rest1 + i_rest1) ...) * ncol + j
// (ld_packed = 0: the packed row length), so several tensors can be packed
// side by side into one wider matrix.
// mode 0 copies src -> packed, mode 1 adds alpha * packed into src.  Every
// element of src maps to exactly one packed element (no write conflicts).
// ---------------------------------------------------------------------------
// Division by a divisor fixed for a launch (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994,
// Thm. 4.2): with l = ceil(log2 d), s = 31 + l and m = ceil(2^s / d),
// floor(x / d) = (x m) >> s exactly for 0 <= x < 2^31 (x m < 2^64).
// The multiplier (< 2^33) and the shift share one word (shift in bits
// 40..47) to keep the batched launch arguments small.
typedef uint64_t FastDiv;
static inline FastDiv fast_div(int64_t d) {
    int l = 0;
    while (((int64_t)1 << l) < d) l++;
    const int s = 31 + l;
    const unsigned __int128 num = ((unsigned __int128)1 << s) + (unsigned __int128)(d - 1);
    return (uint64_t)(num / (unsigned __int128)d) | ((uint64_t)s << 40);
}
__device__ __forceinline__ int64_t fdiv(int64_t x, FastDiv v) {
    return (int64_t)(((uint64_t)x * (v & 0xFFFFFFFFFFull)) >> (int)(v >> 40));
}

struct AxesDesc {
    int nrest;
    int small;                 // every element index < 2^31: the FastDiv forms apply
    int64_t s0, s1;            // strides of the front axes (s1 unused when one)
    int64_t d1;                // dims of the second front axis (1 when one)
    int64_t nfront;            // size of the front index space
    int64_t rest_dim[XR_MAX_FRAG_AX], rest_stride[XR_MAX_FRAG_AX];
    int64_t ncol, xsize, nrows;
    FastDiv fx, fc, fd1, frest[XR_MAX_FRAG_AX];   // by xsize, ncol, d1, rest_dim[k]
};

// offset in the full tensor of packed column x (rest axes and column)
__device__ __forceinline__ int64_t axes_rest_offset(const AxesDesc& ad, int64_t x) {
    const bool sm = ad.small;
    int64_t y = sm ? fdiv(x, ad.fc) : x / ad.ncol;
    int64_t off = x - y * ad.ncol;
    for (int k = ad.nrest - 1; k >= 0; k--) {
        const int64_t q = sm ? fdiv(y, ad.frest[k]) : y / ad.rest_dim[k];
        off += (y - q * ad.rest_dim[k]) * ad.rest_stride[k];
        y = q;
    }
    return off;
}

// Packed element i: its offset in the packed matrix (pi) and in the full
// tensor (off); false when its row maps outside the front index space.
__device__ __forceinline__ bool axes_pack_index(const AxesDesc& ad, const int32_t* __restrict__ rowmap,
                                                int64_t ldp, int64_t i, int64_t* pi, int64_t* off) {
    const bool sm = ad.small;
    const int64_t r = sm ? fdiv(i, ad.fx) : i / ad.xsize, x = i - r * ad.xsize;
    *pi = r * ldp + x;
    const int64_t f = rowmap ? (int64_t)rowmap[r] : r;
    if (f < 0 || f >= ad.nfront) return false;
    const int64_t a = sm ? fdiv(f, ad.fd1) : f / ad.d1, b = f - a * ad.d1;
    *off = a * ad.s0 + b * ad.s1 + axes_rest_offset(ad, x);
    return true;
}

// Grid-stride pass over the packed elements [0, total), XR_PACK_UNROLL per
// thread and round: every element's address is computed and every load
// issued before any store, so a wave keeps that many loads in flight.
#define XR_PACK_UNROLL 4
__device__ __forceinline__ void axes_pack_run(const AxesDesc& ad, const int32_t* __restrict__ rowmap,
                                              double* full, double* packed, int64_t ldp, int mode,
                                              double alpha, int64_t total) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < total; i0 += XR_PACK_UNROLL * stride) {
        int64_t pi[XR_PACK_UNROLL], off[XR_PACK_UNROLL];
        bool ok[XR_PACK_UNROLL];
#pragma unroll
        for (int u = 0; u < XR_PACK_UNROLL; u++) {
            const int64_t i = i0 + u * stride;
            pi[u] = 0;
            off[u] = 0;
            ok[u] = i < total && axes_pack_index(ad, rowmap, ldp, i, &pi[u], &off[u]);
        }
        if (mode == 0) {
            double v[XR_PACK_UNROLL];
#pragma unroll
            for (int u = 0; u < XR_PACK_UNROLL; u++) v[u] = ok[u] ? full[off[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < XR_PACK_UNROLL; u++)
                if (ok[u]) packed[pi[u]] = v[u];
        } else {
            double p[XR_PACK_UNROLL], q[XR_PACK_UNROLL];
#pragma unroll
            for (int u = 0; u < XR_PACK_UNROLL; u++) {
                p[u] = ok[u] ? packed[pi[u]] : 0.0;
                q[u] = ok[u] ? full[off[u]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < XR_PACK_UNROLL; u++)
                if (ok[u]) full[off[u]] = q[u] + alpha * p[u];
        }
    }
}

__global__ void k_axes_pack(AxesDesc ad, const int32_t* __restrict__ rowmap, double* __restrict__ full,
                            double* __restrict__ packed, int64_t ldp, int mode, double alpha) {
    axes_pack_run(ad, rowmap, full, packed, ldp, mode, alpha, ad.nrows * ad.xsize);
}

static int axes_desc(xr_ctx* c, int nf, const int64_t* dims, int64_t ncol, int front0, int front1,
                     const int32_t* d_rowmap, int64_t nrows, AxesDesc* ad) {
    if (nf < 1 || nf > XR_MAX_FRAG_AX) return xr_fail(c, XR_EUNSUPPORTED, "1 to 8 fragment axes");
    if (front0 < 0 || front0 >= nf || front1 >= nf || front1 == front0 || (front1 >= 0 && front1 < front0))
        return xr_fail(c, XR_EINVAL, "front axes must satisfy 0 <= front0 < front1 < nf (front1 = -1 for one)");
    if (ncol <= 0) return xr_fail(c, XR_EINVAL, "bad ncol");
    int64_t stride[XR_MAX_FRAG_AX];
    int64_t acc = ncol;
    for (int m = nf - 1; m >= 0; m--) {
        if (dims[m] <= 0) return xr_fail(c, XR_EINVAL, "empty fragment axis");
        stride[m] = acc;
        acc *= dims[m];
    }
    ad->s0 = stride[front0];
    ad->s1 = front1 >= 0 ? stride[front1] : 0;
    ad->d1 = front1 >= 0 ? dims[front1] : 1;
    ad->nfront = dims[front0] * ad->d1;
    ad->nrest = 0;
    ad->xsize = ncol;
    for (int m = 0; m < nf; m++) {
        if (m == front0 || m == front1) continue;
        ad->rest_dim[ad->nrest] = dims[m];
        ad->rest_stride[ad->nrest] = stride[m];
        ad->nrest++;
        ad->xsize *= dims[m];
    }
    ad->ncol = ncol;
    ad->nrows = nrows;
    if (!d_rowmap && nrows != ad->nfront) return xr_fail(c, XR_ESHAPE, "without a row map nrows must be the front size");
    // element indices (and the front indices, below nfront) under 2^31
    const int64_t lim = (int64_t)1 << 31;
    ad->small = (nrows * ad->xsize < lim && ad->nfront < lim) ? 1 : 0;
    ad->fx = fast_div(ad->xsize);
    ad->fc = fast_div(ncol);
    ad->fd1 = fast_div(ad->d1);
    for (int k = 0; k < ad->nrest; k++) ad->frest[k] = fast_div(ad->rest_dim[k]);
    return XR_OK;
}

static int axes_pack_tiled_one(xr_ctx* c, const AxesDesc& ad, const int32_t* rowmap, double* full,
                               double* packed, int64_t ldp, int mode, double alpha);

extern "C" int xr_axes_pack(xr_ctx* c, int nf, const int64_t* dims, int64_t ncol, int front0, int front1,
                            const int32_t* d_rowmap, int64_t nrows, double* d_full, double* d_packed,
                            int64_t ld_packed, int mode, double alpha) {
    XR_ENTRY(c);
    if (mode != 0 && mode != 1) return xr_fail(c, XR_EINVAL, "bad mode");
    AxesDesc ad;
    int rc = axes_desc(c, nf, dims, ncol, front0, front1, d_rowmap, nrows, &ad);
    if (rc) return rc;
    if (ld_packed == 0) ld_packed = ad.xsize;
    if (ld_packed < ad.xsize) return xr_fail(c, XR_ESHAPE, "ld_packed smaller than the packed row");
    const int64_t total = nrows * ad.xsize;
    if (total <= 0) return XR_OK;
    const int64_t inner = front1 >= 0 ? ad.s1 : ad.s0;
    if (inner == ncol && ad.nrest > 0)          // transposing: LDS tiles (pack_launch below)
        return axes_pack_tiled_one(c, ad, d_rowmap, d_full, d_packed, ld_packed, mode, alpha);
    ProfScope prof(c, XR_K_SCATTER, 0.0, (mode ? 24.0 : 16.0) * total);
    k_axes_pack<<<xr_grid(ceil_div(total, XR_PACK_UNROLL)), 256, 0, c->stream>>>(ad, d_rowmap, d_full, d_packed, ld_packed, mode, alpha);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Several packs (or unpacks) of one mode in one launch: blockIdx.y picks the
// item, the x dimension strides over its elements.  Items must not write
// the same memory (unpacks of one launch target distinct elements).
// When the front axes include the innermost fragment axis (stride ncol),
// consecutive packed columns are far apart in the full tensor while
// consecutive packed rows are adjacent: such an item ("tiled") is moved as
// 64 x 64 tiles through LDS, read along rows on the full side and along
// columns on the packed side, so both sides stay coalesced.
#define XR_PACK_ITEMS 12
struct AxesItem {
    AxesDesc ad;
    const int32_t* rowmap;
    double* full;
    double* packed;
    int64_t ldp;
    int tiled;
};
struct AxesBatch {
    int n, mode;
    double alpha;
    AxesItem it[XR_PACK_ITEMS];
};

__global__ __launch_bounds__(256) void k_axes_pack_multi(AxesBatch bt) {
    const AxesItem& I = bt.it[blockIdx.y];
    axes_pack_run(I.ad, I.rowmap, I.full, I.packed, I.ldp, bt.mode, bt.alpha, I.ad.nrows * I.ad.xsize);
}

__global__ __launch_bounds__(256) void k_axes_pack_tiled(AxesBatch bt) {
    const AxesItem& I = bt.it[blockIdx.y];
    const AxesDesc& ad = I.ad;
    __shared__ double L[64][65];               // L[x - x0][r - r0]
    __shared__ int64_t gx[64];
    __shared__ int rok[64];                    // packed row r0 + i maps to a front index
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t nr = ad.nrows, nx = ad.xsize;
    const int64_t tx = (nx + 63) >> 6, tiles = ((nr + 63) >> 6) * tx;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t r0 = (t / tx) << 6, x0 = (t % tx) << 6;
        const int64_t r = r0 + lane;
        const int64_t f = r < nr ? (I.rowmap ? (int64_t)I.rowmap[r] : r) : -1;
        const bool rv = f >= 0 && f < ad.nfront;    // rows mapped outside are left alone
        const int64_t fr = rv ? (f / ad.d1) * ad.s0 + (f % ad.d1) * ad.s1 : 0;
        __syncthreads();                        // previous tile done with L, gx and rok
        if (tid < 64) {
            gx[tid] = (x0 + tid < nx) ? axes_rest_offset(ad, x0 + tid) : 0;
            rok[tid] = rv ? 1 : 0;
        }
        if (bt.mode == 0) {
            __syncthreads();
            for (int xi = w; xi < 64; xi += 4)
                if (x0 + xi < nx && rv) L[xi][lane] = I.full[fr + gx[xi]];
            __syncthreads();
            for (int ri = w; ri < 64; ri += 4)
                if (r0 + ri < nr && x0 + lane < nx && rok[ri])
                    I.packed[(r0 + ri) * I.ldp + x0 + lane] = L[lane][ri];
        } else {
            for (int ri = w; ri < 64; ri += 4)
                if (r0 + ri < nr && x0 + lane < nx) L[lane][ri] = I.packed[(r0 + ri) * I.ldp + x0 + lane];
            __syncthreads();
            for (int xi = w; xi < 64; xi += 4)
                if (x0 + xi < nx && rv) I.full[fr + gx[xi]] += bt.alpha * L[xi][lane];
        }
    }
}

// Launches the items of one kind (tiled or element-wise), XR_PACK_ITEMS per grid.
static int pack_launch(xr_ctx* c, const std::vector<AxesItem>& items, int tiled, int mode, double alpha) {
    for (size_t base = 0; base < items.size(); base += XR_PACK_ITEMS) {
        AxesBatch bt;
        bt.n = (int)((items.size() - base < XR_PACK_ITEMS) ? items.size() - base : XR_PACK_ITEMS);
        bt.mode = mode;
        bt.alpha = alpha;
        int64_t most = 0, bytes = 0;
        for (int t = 0; t < bt.n; t++) {
            const AxesItem& I = bt.it[t] = items[base + t];
            const int64_t tot = I.ad.nrows * I.ad.xsize;
            const int64_t need = tiled ? ceil_div(I.ad.nrows, 64) * ceil_div(I.ad.xsize, 64)
                                       : (int64_t)xr_grid(ceil_div(tot, XR_PACK_UNROLL));
            most = need > most ? need : most;
            bytes += tot;
        }
        if (bytes <= 0) continue;
        ProfScope prof(c, XR_K_SCATTER, 0.0, (mode ? 24.0 : 16.0) * bytes);
        dim3 grid((unsigned)(most < (1 << 20) ? most : (1 << 20)), (unsigned)bt.n);
        if (tiled) k_axes_pack_tiled<<<grid, 256, 0, c->stream>>>(bt);
        else k_axes_pack_multi<<<grid, 256, 0, c->stream>>>(bt);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

static int axes_pack_tiled_one(xr_ctx* c, const AxesDesc& ad, const int32_t* rowmap, double* full,
                               double* packed, int64_t ldp, int mode, double alpha) {
    AxesItem I;
    I.ad = ad;
    I.rowmap = rowmap;
    I.full = full;
    I.packed = packed;
    I.ldp = ldp;
    I.tiled = 1;
    return pack_launch(c, std::vector<AxesItem>{I}, 1, mode, alpha);
}

extern "C" int xr_axes_pack_multi(xr_ctx* c, int n_items, const int* nf, const int64_t* dims,
                                  const int64_t* ncol, const int* front0, const int* front1,
                                  const int32_t* const* d_rowmaps, const int64_t* nrows,
                                  double* const* d_full, double* const* d_packed,
                                  const int64_t* ld_packed, int mode, double alpha) {
    XR_ENTRY(c);
    if (mode != 0 && mode != 1) return xr_fail(c, XR_EINVAL, "bad mode");
    if (n_items < 0) return xr_fail(c, XR_EINVAL, "negative item count");
    std::vector<AxesItem> plain, tiled;
    for (int q = 0; q < n_items; q++) {
        AxesItem I;
        int rc = axes_desc(c, nf[q], dims + (int64_t)q * XR_MAX_FRAG_AX, ncol[q], front0[q], front1[q],
                           d_rowmaps ? d_rowmaps[q] : nullptr, nrows[q], &I.ad);
        if (rc) return rc;
        I.rowmap = d_rowmaps ? d_rowmaps[q] : nullptr;
        I.full = d_full[q];
        I.packed = d_packed[q];
        I.ldp = ld_packed[q] ? ld_packed[q] : I.ad.xsize;
        if (I.ldp < I.ad.xsize) return xr_fail(c, XR_ESHAPE, "ld_packed smaller than the packed row");
        // innermost front stride: s1 for two front axes, else s0
        const int64_t inner = front1[q] >= 0 ? I.ad.s1 : I.ad.s0;
        I.tiled = (inner == ncol[q] && I.ad.nrest > 0) ? 1 : 0;
        (I.tiled ? tiled : plain).push_back(I);
    }
    int rc = pack_launch(c, plain, 0, mode, alpha);
    if (rc) return rc;
    return pack_launch(c, tiled, 1, mode, alpha);
}

// dst[p, i, q] = sum_j op(U)[i, j] src[p, j, q] (+ beta dst) for many
// (pre, n, post) tensor views in one launch: a small matrix (n <= 64) applied
// along one fragment axis of each item, op(U) = U or U^T.  The columns
// c = p * post + q are cut into tiles of 64; a workgroup stages op(U)^T once
// and then, per tile, the n x 64 operand block in LDS, and wave w computes
// rows [w RB, w RB + RB) of the tile's output, lane = column.  Loads and
// stores run along q (one contiguous run per p); for post == 1 the tile is
// one contiguous 64 x n block, moved through LDS in both directions.
#define XR_AXIS_ITEMS 32
struct AxisItem {
    const double* U;
    const double* src;
    double* dst;
    int64_t pre, post;
    int n, trans;
};
struct AxisBatch {
    int nitems;
    double beta;
    AxisItem it[XR_AXIS_ITEMS];
};

template <int RB>
__global__ __launch_bounds__(256) void k_axis_apply_multi(AxisBatch bt) {
    constexpr int NMAX = 4 * RB;              // largest axis length of this instance
    const AxisItem& I = bt.it[blockIdx.y];
    __shared__ double Ut[NMAX][NMAX];         // Ut[j][i] = op(U)[i][j]
    __shared__ double Xs[NMAX][64 + 1];       // Xs[j][c]: operand tile, then output tile (post == 1)
    const int n = I.n;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t post = I.post, cols = I.pre * post;
    const int64_t tiles = (cols + 63) >> 6;
    const double beta = bt.beta;
    for (int e = tid; e < NMAX * NMAX; e += 256) {
        const int i = e / NMAX, j = e - (e / NMAX) * NMAX;
        double v = 0.0;
        if (i < n && j < n) v = I.trans ? I.U[j * n + i] : I.U[i * n + j];
        Ut[j][i] = v;
    }
    // operand tile t in registers (at most RB values a thread): the next
    // tile's loads are in flight while the current one is computed
    double pf[RB];
    auto fetch = [&](int64_t t) {
        const int64_t c0 = t << 6;
        const int width = (cols - c0 < 64) ? (int)(cols - c0) : 64;
        if (post == 1) {
            const double* src = I.src + c0 * n;
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int e = tid + 256 * u;
                pf[u] = (e < width * n) ? src[e] : 0.0;
            }
        } else {
            const int64_t c = c0 + lane;
            const int64_t p = c / post, q = c - p * post;
            const int64_t base = p * n * post + q;
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int j = w + 4 * u;
                pf[u] = (lane < width && j < n) ? I.src[base + j * post] : 0.0;
            }
        }
    };
    int64_t t = blockIdx.x;
    if (t < tiles) fetch(t);
    for (; t < tiles; t += gridDim.x) {
        const int64_t c0 = t << 6;
        const int width = (cols - c0 < 64) ? (int)(cols - c0) : 64;
        __syncthreads();                      // Ut written / previous tile consumed
        if (post == 1) {
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int e = tid + 256 * u;
                if (e < width * n) {
                    const int c = e / n, j = e - c * n;
                    Xs[j][c] = pf[u];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int j = w + 4 * u;
                if (j < n) Xs[j][lane] = pf[u];
            }
        }
        __syncthreads();
        if (t + gridDim.x < tiles) fetch(t + gridDim.x);
        int64_t base = 0;                     // this lane's column: p * n * post + q
        if (post != 1) {
            const int64_t c = c0 + lane;
            const int64_t p = c / post, q = c - p * post;
            base = p * n * post + q;
        }
        double acc[RB];
#pragma unroll
        for (int k = 0; k < RB; k++) acc[k] = 0.0;
        for (int j = 0; j < n; j++) {
            const double x = Xs[j][lane];
#pragma unroll
            for (int k = 0; k < RB; k++) acc[k] = fma(Ut[j][w * RB + k], x, acc[k]);
        }
        if (post == 1) {
            __syncthreads();                  // every wave has read Xs
#pragma unroll
            for (int k = 0; k < RB; k++)
                if (w * RB + k < n) Xs[w * RB + k][lane] = acc[k];
            __syncthreads();
            double* d = I.dst + c0 * n;
            for (int e = tid; e < width * n; e += 256) {
                const int c = e / n, i = e - c * n;
                const double v = Xs[i][c];
                d[e] = (beta != 0.0) ? v + beta * d[e] : v;
            }
        } else if (lane < width) {
#pragma unroll
            for (int k = 0; k < RB; k++) {
                const int i = w * RB + k;
                if (i < n) {
                    double* d = I.dst + base + (int64_t)i * post;
                    *d = (beta != 0.0) ? acc[k] + beta * *d : acc[k];
                }
            }
        }
    }
}

// Workgroups of 256 threads of an axis-apply instance resident on the
// device at once (CUs x occupancy), per instance.
static int64_t axis_slots(const void* kernel) {
    static std::mutex mu;
    static std::map<const void*, int64_t> memo;
    std::lock_guard<std::mutex> lock(mu);
    auto it = memo.find(kernel);
    if (it != memo.end()) return it->second;
    int dev = 0, cus = 0, per = 0;
    int64_t r = 1536;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) == hipSuccess && cus > 0 && per > 0)
        r = (int64_t)cus * per;
    else
        (void)hipGetLastError();
    memo[kernel] = r;
    return r;
}

extern "C" int xr_axis_apply_multi(xr_ctx* c, int n_items, const double* const* d_U, const int* n,
                                   const int* trans, const int64_t* pre, const int64_t* post,
                                   const double* const* d_src, double* const* d_dst, double beta) {
    XR_ENTRY(c);
    if (n_items < 0) return xr_fail(c, XR_EINVAL, "negative item count");
    for (int base = 0; base < n_items; base += XR_AXIS_ITEMS) {
        AxisBatch bt;
        bt.nitems = (n_items - base < XR_AXIS_ITEMS) ? n_items - base : XR_AXIS_ITEMS;
        bt.beta = beta;
        int64_t most = 0, elems = 0, flops = 0, tiles = 0;
        int nmax = 1;
        for (int t = 0; t < bt.nitems; t++) {
            const int q = base + t;
            if (n[q] < 1 || n[q] > 64) return xr_fail(c, XR_EUNSUPPORTED, "axis length must be 1..64");
            if (pre[q] < 0 || post[q] < 1) return xr_fail(c, XR_EINVAL, "bad pre / post");
            if (d_src[q] == d_dst[q]) return xr_fail(c, XR_EINVAL, "axis apply cannot work in place");
            bt.it[t] = AxisItem{d_U[q], d_src[q], d_dst[q], pre[q], post[q], n[q], trans[q]};
            const int64_t tot = pre[q] * n[q] * post[q];
            most = tot > most ? tot : most;
            elems += tot;
            flops += tot * n[q];
            nmax = n[q] > nmax ? n[q] : nmax;
            const int64_t tq = ceil_div(pre[q] * post[q], 64);
            tiles = tq > tiles ? tq : tiles;
        }
        if (most <= 0) continue;
        ProfScope prof(c, XR_K_GEMM_BW, 2.0 * flops, (beta != 0.0 ? 24.0 : 16.0) * elems);
        // one round of resident workgroups, each looping over at least two of
        // its item's tiles (the next one's loads overlap the current one's
        // work, and op(U) is staged once per workgroup)
        const void* kern = nmax <= 4 ? reinterpret_cast<const void*>(&k_axis_apply_multi<1>)
                         : nmax <= 8 ? reinterpret_cast<const void*>(&k_axis_apply_multi<2>)
                         : nmax <= 16 ? reinterpret_cast<const void*>(&k_axis_apply_multi<4>)
                         : nmax <= 32 ? reinterpret_cast<const void*>(&k_axis_apply_multi<8>)
                         : reinterpret_cast<const void*>(&k_axis_apply_multi<16>);
        int64_t gx = ceil_div(tiles, 2), cap = axis_slots(kern) / bt.nitems;
        if (cap < 1) cap = 1;
        gx = gx < cap ? gx : cap;
        dim3 grid((unsigned)(gx > 0 ? gx : 1), (unsigned)bt.nitems);
        if (nmax <= 4) k_axis_apply_multi<1><<<grid, 256, 0, c->stream>>>(bt);
        else if (nmax <= 8) k_axis_apply_multi<2><<<grid, 256, 0, c->stream>>>(bt);
        else if (nmax <= 16) k_axis_apply_multi<4><<<grid, 256, 0, c->stream>>>(bt);
        else if (nmax <= 32) k_axis_apply_multi<8><<<grid, 256, 0, c->stream>>>(bt);
        else k_axis_apply_multi<16><<<grid, 256, 0, c->stream>>>(bt);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

// A = A + A^T in place for a square n x n matrix: workgroup (a, b) with a <= b
// holds the 32 x 32 tiles (a, b) and (b, a) in LDS and writes both sums (the
// workgroups with a > b have nothing to do).
__global__ __launch_bounds__(256) void k_symmetrize(double* __restrict__ A, int64_t n) {
    const int64_t a = blockIdx.y, b = blockIdx.x;
    if (a > b) return;
    __shared__ double P[32][33], Q[32][33];       // P = tile (a, b), Q = tile (b, a)
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int64_t ra = a * 32, rb = b * 32;
    for (int j = ty; j < 32; j += 8) {
        const int64_t r1 = ra + j, c1 = rb + tx, r2 = rb + j, c2 = ra + tx;
        P[j][tx] = (r1 < n && c1 < n) ? A[r1 * n + c1] : 0.0;
        Q[j][tx] = (r2 < n && c2 < n) ? A[r2 * n + c2] : 0.0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t r1 = ra + j, c1 = rb + tx, r2 = rb + j, c2 = ra + tx;
        if (r1 < n && c1 < n) A[r1 * n + c1] = P[j][tx] + Q[tx][j];
        if (a != b && r2 < n && c2 < n) A[r2 * n + c2] = Q[j][tx] + P[tx][j];
    }
}

extern "C" int xr_symmetrize(xr_ctx* c, double* A, int64_t n) {
    XR_ENTRY(c);
    if (n < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (n == 0) return XR_OK;
    const int64_t t = ceil_div(n, 32);
    if (t > 65535) return xr_fail(c, XR_EUNSUPPORTED, "matrix too large for the symmetrize grid");
    ProfScope prof(c, XR_K_SCATTER, 0.0, 16.0 * n * n);
    k_symmetrize<<<dim3((unsigned)t, (unsigned)t), dim3(32, 8), 0, c->stream>>>(A, n);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// dst[J, I, rev(x)] = src[I, J, x] (or +=) for a density (nI, nJ, d^k): the
// state axes swapped and the index digits reversed (the density of the
// conjugate string between the swapped charges, for real states).  A
// workgroup takes the d x d (first digit, last digit) tile of one (I, J)
// plane for one string of middle digits and turns it through LDS; the tile's
// rows are runs of d contiguous doubles in src and in dst.  k <= 1: plane
// copies.
__global__ __launch_bounds__(256) void k_conjugate(const double* __restrict__ src, int64_t nI, int64_t nJ,
                                                   int d, int k, int64_t P, int64_t nmid,
                                                   double* __restrict__ dst, int acc) {
    __shared__ double T[64][65];
    const int64_t blk = blockIdx.x;
    const int64_t pl = blk / nmid, mid = blk - pl * nmid;
    const int64_t I = pl / nJ, J = pl - I * nJ;
    const double* s = src + pl * P;
    double* o = dst + (J * nI + I) * P;
    if (k <= 1) {
        for (int64_t t = threadIdx.x; t < P; t += blockDim.x) o[t] = acc ? o[t] + s[t] : s[t];
        return;
    }
    int64_t md = 0;
    for (int64_t x = mid, j = 0; j < k - 2; j++, x /= d) md = md * d + x % d;
    const int64_t hi = P / d;
    const int dd = d * d;
    for (int t = threadIdx.x; t < dd; t += blockDim.x) {
        const int p = t / d, q = t - (t / d) * d;
        T[p][q] = s[p * hi + mid * d + q];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < dd; t += blockDim.x) {
        const int q = t / d, p = t - (t / d) * d;
        const int64_t idx = q * hi + md * d + p;
        o[idx] = acc ? o[idx] + T[p][q] : T[p][q];
    }
}

extern "C" int xr_conjugate(xr_ctx* c, const double* src, int64_t nI, int64_t nJ, int d, int k, double* dst,
                            int accumulate) {
    XR_ENTRY(c);
    if (nI < 0 || nJ < 0 || d < 1 || k < 0 || k > 5) return xr_fail(c, XR_EINVAL, "bad density shape");
    if (d > 64) return xr_fail(c, XR_EUNSUPPORTED, "more than 64 spin orbitals");
    if (src == dst && nI * nJ > 0) return xr_fail(c, XR_EINVAL, "xr_conjugate cannot work in place");
    int64_t P = 1;
    for (int j = 0; j < k; j++) P *= d;
    int64_t nmid = 1;
    for (int j = 2; j < k; j++) nmid *= d;
    const int64_t blocks = nI * nJ * nmid;
    if (blocks == 0) return XR_OK;
    if (blocks > 0x7fffffffLL) return xr_fail(c, XR_EUNSUPPORTED, "density too large for the conjugate grid");
    ProfScope prof(c, XR_K_SCATTER, 0.0, (accumulate ? 24.0 : 16.0) * nI * nJ * P);
    k_conjugate<<<(unsigned)blocks, 256, 0, c->stream>>>(src, nI, nJ, d, k, P, nmid, dst, accumulate ? 1 : 0);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Pair-block adjoint input from a reduced pair matrix G2 (rows and columns
// over the pair index a * d2 + b, reordered through inv when given):
//   out[(I1*nJ1+J1)*nB + I2*nJ2+J2] = G2[row(r1[I1], r2[I2]), row(c1[J1], c2[J2])]
// (out[b*nA + a] when swapped), row(a, b) = (inv ? inv[a*d2 + b] : a*d2 + b) - base
// (base: first row of the block of G2 that is passed).
__global__ void k_pair_gather(const double* __restrict__ G2, int64_t ldg, int64_t d2, int64_t base,
                              const int32_t* __restrict__ inv, const int64_t* __restrict__ r1,
                              const int64_t* __restrict__ c1, const int64_t* __restrict__ r2,
                              const int64_t* __restrict__ c2, int64_t nI1, int64_t nJ1, int64_t nI2,
                              int64_t nJ2, int swapped, double* __restrict__ out) {
    const int64_t nA = nI1 * nJ1, nB = nI2 * nJ2;
    XR_GRID_STRIDE(i, nA * nB) {
        int64_t a, b;
        if (swapped) { b = i / nA; a = i - b * nA; }
        else { a = i / nB; b = i - a * nB; }
        const int64_t I1 = a / nJ1, J1 = a - I1 * nJ1, I2 = b / nJ2, J2 = b - I2 * nJ2;
        int64_t row = r1[I1] * d2 + r2[I2], col = c1[J1] * d2 + c2[J2];
        if (inv) { row = inv[row]; col = inv[col]; }
        out[i] = G2[(row - base) * ldg + (col - base)];
    }
}

extern "C" int xr_pair_gather(xr_ctx* c, const double* d_G2, int64_t ldg, int64_t d2, int64_t base,
                              const int32_t* d_inv, const int64_t* d_r1, const int64_t* d_c1,
                              const int64_t* d_r2, const int64_t* d_c2, int64_t nI1, int64_t nJ1,
                              int64_t nI2, int64_t nJ2, int swapped, double* d_out) {
    XR_ENTRY(c);
    const int64_t n = nI1 * nJ1 * nI2 * nJ2;
    if (n <= 0) return XR_OK;
    ProfScope prof(c, XR_K_SCATTER, 0.0, 16.0 * n);
    k_pair_gather<<<xr_grid(n), 256, 0, c->stream>>>(d_G2, ldg, d2, base, d_inv, d_r1, d_c1, d_r2, d_c2,
                                                     nI1, nJ1, nI2, nJ2, swapped, d_out);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

struct BraketDesc {
    int nf;
    int64_t spf;
    int64_t pw[XR_MAX_FRAG];  // spf^(nf-1-m)
};

// One thread per element of the tensor-product matrix: the sums of
// general-XRCC/hamiltonian.py:21-63 (monomer terms, pair couplings, trimer
// couplings by number of fragments whose state changes).  With par, one-
// electron transfers between non-adjacent fragments carry the sign of the
// electron counts of the spectator fragments between them.
__device__ __forceinline__ void k_braket_elem(int64_t idx, BraketDesc bd, const double* const* __restrict__ H1,
                         const double* const* __restrict__ H2, const double* const* __restrict__ H3,
                         const int8_t* __restrict__ par, double* __restrict__ out, int64_t n) {
    const int nf = bd.nf;
    const int64_t s = bd.spf;
    int64_t I = idx / n, J = idx - (idx / n) * n;
    int b[XR_MAX_FRAG], k[XR_MAX_FRAG];
    int tr[3];
    int nt = 0;
    for (int m = 0; m < nf; m++) {
        b[m] = (int)((I / bd.pw[m]) % s);
        k[m] = (int)((J / bd.pw[m]) % s);
        if (b[m] != k[m]) {
            if (nt < 3) tr[nt] = m;
            nt++;
        }
    }
    double v = 0.0;
    auto h2 = [&](int M, int N, int64_t r, int64_t c_) { return H2[M * nf + N][r * s * s + c_]; };
    auto h3 = [&](int M, int N, int O, int64_t r, int64_t c_) {
        const double* p = H3 ? H3[(M * nf + N) * nf + O] : nullptr;
        return p ? p[r * s * s * s + c_] : 0.0;
    };
    if (nt == 0) {
        for (int M = 0; M < nf; M++) {
            v += H1[M][b[M] * s + k[M]];
            for (int N = M + 1; N < nf; N++) v += h2(M, N, b[M] * s + b[N], k[M] * s + k[N]);
        }
    } else if (nt == 1) {
        int M = tr[0];
        v += H1[M][b[M] * s + k[M]];
        for (int N = 0; N < M; N++) v += h2(N, M, b[N] * s + b[M], k[N] * s + k[M]);
        for (int N = M + 1; N < nf; N++) v += h2(M, N, b[M] * s + b[N], k[M] * s + k[N]);
    } else if (nt == 2) {
        int M = tr[0], N = tr[1];
        double sg = 1.0;
        if (par && par[M * s + b[M]] != par[M * s + k[M]])
            for (int O = M + 1; O < N; O++)
                if (par[O * s + b[O]]) sg = -sg;
        v += sg * h2(M, N, b[M] * s + b[N], k[M] * s + k[N]);
        for (int O = 0; O < M; O++)
            v += h3(O, M, N, (b[O] * s + b[M]) * s + b[N], (k[O] * s + k[M]) * s + k[N]);
        for (int O = M + 1; O < N; O++)
            v += h3(M, O, N, (b[M] * s + b[O]) * s + b[N], (k[M] * s + k[O]) * s + k[N]);
        for (int O = N + 1; O < nf; O++)
            v += h3(M, N, O, (b[M] * s + b[N]) * s + b[O], (k[M] * s + k[N]) * s + k[O]);
    } else if (nt == 3) {
        int M = tr[0], N = tr[1], O = tr[2];
        v += h3(M, N, O, (b[M] * s + b[N]) * s + b[O], (k[M] * s + k[N]) * s + k[O]);
    }
    out[idx] = v;
}

__global__ void k_braket(BraketDesc bd, const double* const* __restrict__ H1,
                         const double* const* __restrict__ H2, const double* const* __restrict__ H3,
                         const int8_t* __restrict__ par, double* __restrict__ out, int64_t n) {
    XR_GRID_STRIDE(idx, n * n) k_braket_elem(idx, bd, H1, H2, H3, par, out, n);
}

extern "C" int xr_braket_assemble(xr_ctx* c, int n_frag, int64_t spf, const double* const* H1,
                                  const double* const* H2, const double* const* H3,
                                  const int8_t* d_par, double* d_out) {
    XR_ENTRY(c);
    if (n_frag < 1 || n_frag > XR_MAX_FRAG) return xr_fail(c, XR_EUNSUPPORTED, "1 to 8 fragments");
    BraketDesc bd;
    bd.nf = n_frag;
    bd.spf = spf;
    int64_t n = 1;
    for (int m = n_frag - 1; m >= 0; m--) {
        bd.pw[m] = n;
        n *= spf;
    }
    if (n > 3037000499LL) return xr_fail(c, XR_EUNSUPPORTED, "matrix too large");
    // device copies of the pointer tables
    const int np = n_frag + n_frag * n_frag + (H3 ? n_frag * n_frag * n_frag : 0);
    const double** tbl = (const double**)xr_workspace(c, 30, np * sizeof(double*));
    if (!tbl) return xr_fail(c, XR_ENOMEM, "workspace");
    std::vector<const double*> h(np, nullptr);
    for (int m = 0; m < n_frag; m++) h[m] = H1[m];
    for (int m = 0; m < n_frag * n_frag; m++) h[n_frag + m] = H2[m];
    if (H3)
        for (int m = 0; m < n_frag * n_frag * n_frag; m++) h[n_frag + n_frag * n_frag + m] = H3[m];
    XR_HIP(c, hipMemcpyAsync(tbl, h.data(), np * sizeof(double*), hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    ProfScope prof(c, XR_K_BRAKET, 0.0, 8.0 * n * n);
    k_braket<<<xr_grid(n * n), 256, 0, c->stream>>>(
        bd, tbl, tbl + n_frag, H3 ? tbl + n_frag + n_frag * n_frag : nullptr, d_par, d_out, n);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Adjoint of the pair-block scatter for a rank-1 XR derivative v v^T on a
// dimer: G[(I1*nJ1+J1), (I2*nJ2+J2)] = v[r1[I1]*d2 + r2[I2]] * v[c1[J1]*d2 + c2[J2]]
// (transposed layout when swapped).
__device__ __forceinline__ void k_pair_outer_elem(int64_t i, const double* __restrict__ v, int64_t d2, const int64_t* __restrict__ r1,
                             const int64_t* __restrict__ c1, const int64_t* __restrict__ r2,
                             const int64_t* __restrict__ c2, int64_t nI1, int64_t nJ1, int64_t nI2,
                             int64_t nJ2, int swapped, double* __restrict__ G) {
    int64_t nA = nI1 * nJ1, nB = nI2 * nJ2;
    int64_t a, b;
    if (swapped) { b = i / nA; a = i - b * nA; }
    else { a = i / nB; b = i - a * nB; }
    int64_t I1 = a / nJ1, J1 = a - I1 * nJ1, I2 = b / nJ2, J2 = b - I2 * nJ2;
    G[i] = v[r1[I1] * d2 + r2[I2]] * v[c1[J1] * d2 + c2[J2]];
}

__global__ void k_pair_outer(const double* __restrict__ v, int64_t d2, const int64_t* __restrict__ r1,
                             const int64_t* __restrict__ c1, const int64_t* __restrict__ r2,
                             const int64_t* __restrict__ c2, int64_t nI1, int64_t nJ1, int64_t nI2,
                             int64_t nJ2, int swapped, double* __restrict__ G) {
    int64_t nA = nI1 * nJ1, nB = nI2 * nJ2;
    XR_GRID_STRIDE(i, nA * nB) k_pair_outer_elem(i, v, d2, r1, c1, r2, c2, nI1, nJ1, nI2, nJ2, swapped, G);
}

extern "C" int xr_pair_outer(xr_ctx* c, const double* d_v, int64_t d2, const int64_t* d_r1,
                             const int64_t* d_c1, const int64_t* d_r2, const int64_t* d_c2,
                             int64_t nI1, int64_t nJ1, int64_t nI2, int64_t nJ2, int swapped,
                             double* d_G) {
    XR_ENTRY(c);
    int64_t n = nI1 * nJ1 * nI2 * nJ2;
    if (n <= 0) return XR_OK;
    ProfScope prof(c, XR_K_SCATTER, 0.0, 8.0 * n);
    k_pair_outer<<<xr_grid(n), 256, 0, c->stream>>>(d_v, d2, d_r1, d_c1, d_r2, d_c2, nI1, nJ1,
                                                          nI2, nJ2, swapped, d_G);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}
