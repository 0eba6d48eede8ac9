// Certified lower bounds for the lowest eigenvalue of dense symmetric
// matrices, the ingredients of the charge sectors' pair-decomposition bound
// of the XR operator: the dressed fixed-charge pair blocks, a batched Lanczos
// run that estimates each block's lowest eigenvalue (and bounds its infinity
// norm), and the diagonal shift A - t I that a Cholesky factorisation then
// tests for positive definiteness.
// ---------------------------------------------------------------------------
#define XR_LZ_MAXB 64
enum { WS_LZ_VEC = 91 };

// out[r, c] = (B[r, c] + B[c, r]) / 2
//           + wm [b_r == b_c] (hm[a_r, a_c] + hm[a_c, a_r]) / 2
//           + wn [a_r == a_c] (hn[b_r, b_c] + hn[b_c, b_r]) / 2
// for the cnt x cnt block B over pair states (a_r, b_r): the symmetric part of
// the pair block plus the shares wm, wn of the two monomer terms.  One 32 x 32
// output tile per workgroup; the transposed input tile goes through LDS.
__global__ __launch_bounds__(256) void k_pair_dress(int64_t cnt, const double* __restrict__ B, int64_t ldb,
                                                    const int* __restrict__ pa, const int* __restrict__ pb,
                                                    const double* __restrict__ hm, int64_t ldm,
                                                    const double* __restrict__ hn, int64_t ldn, double wm,
                                                    double wn, double* __restrict__ out, int64_t ldo) {
    __shared__ double T[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {          // T[j][i] = B[c0 + j, r0 + i]
        const int64_t r = c0 + j, cc = r0 + tx;
        T[j][tx] = (r < cnt && cc < cnt) ? B[r * ldb + cc] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int64_t r = r0 + i, cc = c0 + tx;
        if (r < cnt && cc < cnt) {
            double v = 0.5 * (B[r * ldb + cc] + T[tx][i]);
            const int64_t ar = pa[r], br = pb[r], ac = pa[cc], bc = pb[cc];
            if (br == bc) v += wm * (0.5 * (hm[ar * ldm + ac] + hm[ac * ldm + ar]));
            if (ar == ac) v += wn * (0.5 * (hn[br * ldn + bc] + hn[bc * ldn + br]));
            out[r * ldo + cc] = v;
        }
    }
}

// A batch of up to 64 symmetric matrices for the Lanczos passes.  Matrix b
// is n[b] x n[b] with row stride ld[b]; its vectors sit at voff[b] in the
// workspace: v, the previous v, w = A v and the absolute row sums (n[b]
// each).  The matrix-vector pass gives each matrix ceil(n / 4) workgroups of
// 4 waves (one row per wave), starting at workgroup wg0[b].
struct LzBatch {
    int nb;
    int64_t n[XR_LZ_MAXB];
    int64_t ld[XR_LZ_MAXB];
    int64_t voff[XR_LZ_MAXB];
    int64_t wg0[XR_LZ_MAXB + 1];
    const double* M[XR_LZ_MAXB];
};

// Deterministic start components in [-1, 1): a 64-bit mix of (matrix, row).
__device__ __forceinline__ double lz_start(int b, int64_t i) {
    uint64_t z = ((uint64_t)(uint32_t)b << 40) ^ (uint64_t)i ^ 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__device__ __forceinline__ double block_max256(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// v = the normalised start vector, previous v = 0 (one workgroup per matrix).
__global__ __launch_bounds__(256) void k_lz_init(LzBatch bt, double* V) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    const int64_t n = bt.n[b];
    double* v = V + bt.voff[b];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double x = lz_start(b, i);
        v[i] = x;
        v[n + i] = 0.0;
        s += x * x;
    }
    const double inv = 1.0 / sqrt(block_sum256(s, red));
    for (int64_t i = threadIdx.x; i < n; i += 256) v[i] *= inv;
}

// w = A v for every matrix of the batch, one row per wave (coalesced row
// reads, butterfly reduction); with rowsum set, also the row's absolute sum.
__global__ __launch_bounds__(256) void k_lz_matvec(LzBatch bt, double* V, int rowsum) {
    const int64_t g = blockIdx.x;
    int b = 0;
    while (b + 1 < bt.nb && bt.wg0[b + 1] <= g) b++;
    const int64_t n = bt.n[b];
    const int64_t r = (g - bt.wg0[b]) * 4 + (threadIdx.x >> 6);
    if (r >= n) return;                         // uniform over the wave: no barrier follows
    const int lane = threadIdx.x & 63;
    const double* row = bt.M[b] + r * bt.ld[b];
    const double* v = V + bt.voff[b];
    double acc = 0.0, ab = 0.0;
#pragma unroll 4
    for (int64_t j = lane; j < n; j += 64) {
        const double x = row[j];
        acc += x * v[j];
        ab += fabs(x);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o);
        ab += __shfl_xor(ab, o);
    }
    if (lane == 0) {
        V[bt.voff[b] + 2 * n + r] = acc;
        if (rowsum) V[bt.voff[b] + 3 * n + r] = ab;
    }
}

// One Lanczos step per matrix: alpha = v.w, w -= alpha v + beta_prev v_prev,
// beta = |w|, (v_prev, v) = (v, w / beta) (v = 0 after a breakdown, beta = 0);
// alpha[b][k], beta[b][k] written, and at k = 0 anorm[b] = the largest
// absolute row sum (an upper bound of the 2-norm).  Fixed-order sums.
__global__ __launch_bounds__(256) void k_lz_update(LzBatch bt, double* V, int k, int steps,
                                                   double* __restrict__ alpha, double* __restrict__ beta,
                                                   double* __restrict__ anorm) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    const int64_t n = bt.n[b];
    double* v = V + bt.voff[b];
    double* vp = v + n;
    double* w = v + 2 * n;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += v[i] * w[i];
    const double al = block_sum256(s, red);
    const double bp = k > 0 ? beta[(int64_t)b * steps + k - 1] : 0.0;
    double q = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double x = w[i] - al * v[i] - bp * vp[i];
        w[i] = x;
        q += x * x;
    }
    const double be = sqrt(block_sum256(q, red));
    const double inv = be > 0.0 ? 1.0 / be : 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        vp[i] = v[i];
        v[i] = w[i] * inv;
    }
    if (k == 0) {
        const double* rs = v + 3 * n;
        double mx = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += 256) mx = fmax(mx, rs[i]);
        mx = block_max256(mx, red);
        if (threadIdx.x == 0) anorm[b] = mx;
    }
    if (threadIdx.x == 0) {
        alpha[(int64_t)b * steps + k] = al;
        beta[(int64_t)b * steps + k] = be;
    }
}

// A[b][i, i] += s[b] for up to 64 square matrices (one grid-stride pass over
// all their diagonals; off[b] = the first diagonal element of matrix b).
struct DiagBatch {
    int nb;
    int64_t n[XR_LZ_MAXB];
    int64_t ld[XR_LZ_MAXB];
    int64_t off[XR_LZ_MAXB + 1];
    double* M[XR_LZ_MAXB];
    double s[XR_LZ_MAXB];
};

__global__ __launch_bounds__(256) void k_add_diagonal(DiagBatch db) {
    XR_GRID_STRIDE(e, db.off[db.nb]) {
        int b = 0;
        while (b + 1 < db.nb && db.off[b + 1] <= e) b++;
        const int64_t i = e - db.off[b];
        db.M[b][i * db.ld[b] + i] += db.s[b];
    }
}

extern "C" int xr_pair_dress(xr_ctx* c, int64_t cnt, const double* B, int64_t ldb, const int* pa, const int* pb,
                             const double* hm, int64_t ldm, const double* hn, int64_t ldn, double wm, double wn,
                             double* out, int64_t ldo) {
    XR_ENTRY(c);
    if (cnt < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (cnt == 0) return XR_OK;
    if (ldb < cnt || ldo < cnt || ldm < 1 || ldn < 1) return xr_fail(c, XR_ESHAPE, "leading dimension too small");
    if (!B || !pa || !pb || !hm || !hn || !out) return xr_fail(c, XR_EINVAL, "null operand");
    const int64_t t = ceil_div(cnt, 32);
    if (t > 65535) return xr_fail(c, XR_EUNSUPPORTED, "block too large for the dress grid");
    ProfScope prof(c, XR_K_OTHER, 0.0, 24.0 * cnt * cnt);
    k_pair_dress<<<dim3((unsigned)t, (unsigned)t), 256, 0, c->stream>>>(cnt, B, ldb, pa, pb, hm, ldm, hn, ldn, wm,
                                                                        wn, out, ldo);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

extern "C" int xr_lanczos(xr_ctx* c, int nb, const int64_t* n, const double* const* M, const int64_t* ld,
                          int steps, double* alpha, double* beta, double* anorm) {
    XR_ENTRY(c);
    if (nb < 0 || nb > XR_LZ_MAXB) return xr_fail(c, XR_EINVAL, "between 0 and 64 matrices per call");
    if (steps < 1) return xr_fail(c, XR_EINVAL, "at least one Lanczos step");
    if (nb == 0) return XR_OK;
    if (!n || !M || !ld || !alpha || !beta || !anorm) return xr_fail(c, XR_EINVAL, "null argument");
    LzBatch bt{};
    bt.nb = nb;
    int64_t off = 0, wg = 0;
    double bytes = 0.0;
    for (int b = 0; b < nb; b++) {
        if (n[b] < 1) return xr_fail(c, XR_EINVAL, "empty matrix");
        if (ld[b] < n[b]) return xr_fail(c, XR_ESHAPE, "row stride below the matrix size");
        if (!M[b]) return xr_fail(c, XR_EINVAL, "null matrix");
        bt.n[b] = n[b];
        bt.ld[b] = ld[b];
        bt.M[b] = M[b];
        bt.voff[b] = off;
        bt.wg0[b] = wg;
        off += 4 * n[b];
        wg += ceil_div(n[b], 4);
        bytes += 8.0 * n[b] * n[b];
    }
    bt.wg0[nb] = wg;
    if (wg > 0x7fffffff) return xr_fail(c, XR_EUNSUPPORTED, "matrix-vector grid too large");
    double* V = (double*)xr_workspace(c, WS_LZ_VEC, (size_t)off * sizeof(double));
    if (!V) return xr_fail(c, XR_ENOMEM, "workspace (Lanczos vectors)");
    k_lz_init<<<nb, 256, 0, c->stream>>>(bt, V);
    XR_LAUNCH_CHECK(c);
    for (int k = 0; k < steps; k++) {
        {
            ProfScope prof(c, XR_K_OTHER, 0.0, bytes);
            k_lz_matvec<<<(unsigned)wg, 256, 0, c->stream>>>(bt, V, k == 0);
            XR_LAUNCH_CHECK(c);
        }
        k_lz_update<<<nb, 256, 0, c->stream>>>(bt, V, k, steps, alpha, beta, anorm);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

extern "C" int xr_add_diagonal(xr_ctx* c, int nb, const int64_t* n, double* const* M, const int64_t* ld,
                               const double* shift) {
    XR_ENTRY(c);
    if (nb < 0 || nb > XR_LZ_MAXB) return xr_fail(c, XR_EINVAL, "between 0 and 64 matrices per call");
    if (nb == 0) return XR_OK;
    if (!n || !M || !ld || !shift) return xr_fail(c, XR_EINVAL, "null argument");
    DiagBatch db{};
    db.nb = nb;
    int64_t off = 0;
    for (int b = 0; b < nb; b++) {
        if (n[b] < 0) return xr_fail(c, XR_EINVAL, "negative size");
        if (n[b] > 0 && (ld[b] < n[b] || !M[b])) return xr_fail(c, XR_ESHAPE, "bad matrix or row stride");
        db.n[b] = n[b];
        db.ld[b] = ld[b];
        db.M[b] = M[b];
        db.s[b] = shift[b];
        db.off[b] = off;
        off += n[b];
    }
    db.off[nb] = off;
    if (off == 0) return XR_OK;
    k_add_diagonal<<<xr_grid(off), 256, 0, c->stream>>>(db);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// ---------------------------------------------------------------------------
// Positive-definiteness test by Cholesky factorisation, in place, of up to 64
// matrices at once: right-looking with 64-column panels.  Step p:
//   k_chol_diag    one wave per matrix factors the diagonal block
//                  L11 L11^T = A11 and forms L11^-1;
//   panel GEMM     W = A21 L11^-T (grouped, into a workspace);
//   trailing GEMM  A22 -= W W^T on the lower-triangle tiles (grouped,
//                  accumulating).
// info[b] = 0 when matrix b factored, else the 1-based column whose pivot was
// not positive (NaN included).  Only the outcome is kept; the matrices are
// left overwritten.
// ---------------------------------------------------------------------------
enum { WS_CHOL = 92 };

struct CholBatch {
    int nb;
    int p;                                  // panel index
    int64_t n[XR_LZ_MAXB];
    int64_t ld[XR_LZ_MAXB];
    double* M[XR_LZ_MAXB];
    int id[XR_LZ_MAXB];                     // the listed matrix's entry in info and in the L11^-1 slots
};

__global__ void k_zero_int(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// A double from lane `src` of the wave (src uniform): two scalar reads.
__device__ __forceinline__ double lane_read(double x, int src) {
    const unsigned long long v = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One 64-lane wave per listed matrix.  Lane r keeps row r of the diagonal
// block (lower triangle, identity-padded beyond n) in registers; column j of
// L is computed in every lane and its entries reach the other lanes by
// v_readlane, so the 64 column steps need no barrier.  L then goes to LDS and
// lane j solves L x = e_j for column j of L11^-1, written row-major (zeros
// above the diagonal) to Linv[id].
__global__ __launch_bounds__(64) void k_chol_diag(CholBatch cb, int* __restrict__ info, double* __restrict__ Linv) {
    __shared__ double Ls[64][65];
    const int b = blockIdx.x;
    const int id = cb.id[b];
    if (info[id] != 0) return;              // failed at an earlier step (uniform)
    const int64_t n = cb.n[b], ld = cb.ld[b];
    const double* A = cb.M[b];
    const int64_t c0 = (int64_t)cb.p * 64;
    const int nk = (int)(n - c0 < 64 ? n - c0 : 64);
    const int lane = threadIdx.x;
    double a[64];
#pragma unroll
    for (int k = 0; k < 64; k++)
        a[k] = (lane < nk && k < nk) ? (k <= lane ? A[(c0 + lane) * ld + c0 + k] : 0.0) : (lane == k ? 1.0 : 0.0);
    int fail = 0;
#pragma unroll
    for (int j = 0; j < 64; j++) {
        if (!fail) {
            const double d = lane_read(a[j], j);
            if (!(d > 0.0)) {
                fail = j + 1;               // d is the same in every lane: uniform
            } else {
                const double s = sqrt(d);
                const double rs = 1.0 / s;
                const double l = lane > j ? a[j] * rs : (lane == j ? s : 0.0);
                a[j] = l;
#pragma unroll
                for (int c = j + 1; c < 64; c++) {
                    const double lc = lane_read(l, c);
                    if (lane >= c) a[c] -= l * lc;
                }
            }
        }
    }
    if (fail) {
        if (lane == 0) info[id] = (int)(c0 + fail);
        return;
    }
#pragma unroll
    for (int k = 0; k < 64; k++) Ls[lane][k] = k <= lane ? a[k] : 0.0;
    __syncthreads();
    // lane j: x_r = (delta_rj - sum_{k<r} L[r][k] x_k) / L[r][r], x_r = 0 for r < j
    double* out = Linv + (int64_t)id * 4096;
    double x[64];
#pragma unroll
    for (int r = 0; r < 64; r++) {
        double acc = lane == r ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < r; k++) acc -= Ls[r][k] * x[k];
        x[r] = lane <= r ? acc / Ls[r][r] : 0.0;
        out[r * 64 + lane] = x[r];
    }
}

extern "C" int xr_cholesky_test(xr_ctx* c, int nb, const int64_t* n, double* const* M, const int64_t* ld,
                                int* info) {
    XR_ENTRY(c);
    if (nb < 0 || nb > XR_LZ_MAXB) return xr_fail(c, XR_EINVAL, "between 0 and 64 matrices per call");
    if (nb == 0) return XR_OK;
    if (!n || !M || !ld || !info) return xr_fail(c, XR_EINVAL, "null argument");
    int64_t nmax = 0, wrows = 0;
    std::vector<int64_t> woff(nb);
    for (int b = 0; b < nb; b++) {
        if (n[b] < 1) return xr_fail(c, XR_EINVAL, "empty matrix");
        if (ld[b] < n[b]) return xr_fail(c, XR_ESHAPE, "row stride below the matrix size");
        if (!M[b]) return xr_fail(c, XR_EINVAL, "null matrix");
        nmax = std::max(nmax, n[b]);
        woff[b] = wrows;
        wrows += n[b];
    }
    if (nmax > (int64_t)1 << 30) return xr_fail(c, XR_EUNSUPPORTED, "matrix too large");
    // workspace: L11^-1 of every matrix (64 x 64), then W (n[b] x 64 rows each)
    double* ws = (double*)xr_workspace(c, WS_CHOL, sizeof(double) * (size_t)(4096 * nb + 64 * wrows));
    if (!ws) return xr_fail(c, XR_ENOMEM, "workspace (Cholesky panels)");
    double* Linv = ws;
    double* W = ws + 4096 * (int64_t)nb;
    k_zero_int<<<1, 64, 0, c->stream>>>(info, nb);
    XR_LAUNCH_CHECK(c);
    std::vector<int64_t> Ms, Ns, Ks, lda, ldb, ldc;
    std::vector<const double*> As, Bs;
    std::vector<double*> Cs;
    std::vector<int> wide, narrow;
    for (int64_t p = 0; p * 64 < nmax; p++) {
        CholBatch cb{};
        cb.p = (int)p;
        int k = 0;
        for (int b = 0; b < nb; b++) {
            if (n[b] <= p * 64) continue;   // already factored
            cb.n[k] = n[b];
            cb.ld[k] = ld[b];
            cb.M[k] = M[b];
            cb.id[k] = b;
            k++;
        }
        cb.nb = k;
        {
            ProfScope prof(c, XR_K_OTHER, 0.0, 0.0);
            k_chol_diag<<<(unsigned)k, 64, 0, c->stream>>>(cb, info, Linv);
            XR_LAUNCH_CHECK(c);
        }
        // W = A21 L11^-T for every matrix with rows below the panel
        Ms.clear(); Ns.clear(); Ks.clear(); lda.clear(); ldb.clear(); ldc.clear();
        As.clear(); Bs.clear(); Cs.clear(); wide.clear(); narrow.clear();
        for (int q = 0; q < k; q++) {
            const int64_t rest = cb.n[q] - (p + 1) * 64;
            if (rest <= 0) continue;
            const int b = cb.id[q];
            Ms.push_back(rest);
            Ns.push_back(64);
            Ks.push_back(64);
            As.push_back(cb.M[q] + ((p + 1) * 64) * cb.ld[q] + p * 64);
            lda.push_back(cb.ld[q]);
            Bs.push_back(Linv + 4096 * (int64_t)b);
            ldb.push_back(64);
            Cs.push_back(W + 64 * woff[b]);
            ldc.push_back(64);
        }
        for (size_t q = 0; q < Ms.size(); q++) {
            if (Ms[q] > 64) {
                narrow.push_back((int)q);
            } else {
                const int rc = xr_dgemm_launch(c, 0, 1, Ms[q], 64, 64, 1.0, As[q], lda[q], 0, Bs[q], 64, 0, 0.0,
                                               Cs[q], 64, 0, 1);
                if (rc) return rc;
            }
        }
        if (!narrow.empty()) {
            const int rc = launch_grouped<128, 64>(c, narrow, 0, 1, Ms.data(), Ns.data(), Ks.data(), 1.0, As.data(),
                                                   lda.data(), Bs.data(), ldb.data(), Cs.data(), ldc.data(), 0.0, 0,
                                                   XR_K_OTHER);
            if (rc) return rc;
        }
        // A22 -= W W^T (lower-triangle tiles)
        std::vector<const double*> Ws(Cs.begin(), Cs.end());
        std::vector<double*> Cu;
        std::vector<int64_t> ldu;
        for (size_t q = 0, r = 0; q < (size_t)k; q++) {
            const int64_t rest = cb.n[q] - (p + 1) * 64;
            if (rest <= 0) continue;
            Cu.push_back(cb.M[q] + ((p + 1) * 64) * cb.ld[q] + (p + 1) * 64);
            ldu.push_back(cb.ld[q]);
            if (Ms[r] > 64) wide.push_back((int)r);
            else {
                const int rc = xr_dgemm_launch(c, 0, 1, Ms[r], Ms[r], 64, -1.0, Ws[r], 64, 0, Ws[r], 64, 0, 1.0,
                                               Cu.back(), ldu.back(), 0, 1);
                if (rc) return rc;
            }
            r++;
        }
        if (!wide.empty()) {
            const int rc = launch_grouped<128, 128>(c, wide, 0, 1, Ms.data(), Ms.data(), Ks.data(), -1.0, Ws.data(),
                                                    ldc.data(), Ws.data(), ldc.data(), Cu.data(), ldu.data(), 1.0, 1,
                                                    XR_K_OTHER);
            if (rc) return rc;
        }
    }
    return XR_OK;
}
