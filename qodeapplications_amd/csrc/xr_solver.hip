// Vector kernels of the block-Davidson eigensolver (the matrix-free XR
// solve and the dense fallback): the diagonal preconditioner's divide, the
// residual block with its column norms, column sums of squares, and the
// transposed store of a block of vectors into the basis kept one vector per
// row.  Each is one pass over the n x b block; the column reductions go
// through fixed-order partial sums (the same result on every run).
// ---------------------------------------------------------------------------
#define XR_SOLVER_MAXB 32
enum { WS_SOLVER_PART = 90 };

// Y[i, j] = X[i, j] / d with d = s (lam[i] - theta[j]), replaced by +floor
// where |d| < floor.  X and Y may be the same array.
__global__ void k_precond_divide(int64_t n, int64_t b, const double* X, int64_t ldx, double* Y, int64_t ldy,
                                 const double* __restrict__ lam, const double* __restrict__ theta, double s,
                                 double floor_) {
    XR_GRID_STRIDE(e, n * b) {
        const int64_t i = e / b, j = e - i * b;
        double d = s * (lam[i] - theta[j]);
        if (fabs(d) < floor_) d = floor_;
        Y[i * ldy + j] = X[i * ldx + j] / d;
    }
}

// Fixed-order sum of one value per thread over the block (256 threads).
__device__ __forceinline__ double block_sum256(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// R = AX - X diag(theta) (row-major n x NB blocks, row strides ldx / ldax /
// ldr; R may be null), and part[block][j] = the block's sum of R[i][j]^2 for
// j < k (R[i][j] = A[i][j] when X is null: plain column sums of squares).
template <int NB>
__global__ __launch_bounds__(256) void k_residual_part(int64_t n, int b, int k, const double* X, int64_t ldx,
                                                       const double* A, int64_t lda,
                                                       const double* __restrict__ theta, double* R, int64_t ldr,
                                                       double* __restrict__ part) {
    __shared__ double red[256];
    double acc[NB];
    double th[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        acc[j] = 0.0;
        th[j] = (X && j < b) ? theta[j] : 0.0;
    }
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < b) {
                double r = A[i * lda + j];
                if (X) r -= X[i * ldx + j] * th[j];
                if (R) R[i * ldr + j] = r;
                if (j < k) acc[j] += r * r;
            }
        }
    }
    for (int j = 0; j < k; j++) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if (q == j) v = acc[q];
        const double s = block_sum256(v, red);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * k + j] = s;
    }
}

// out[j] = sum over blocks of part[block][j], or its root: one workgroup per
// column, each thread summing every 256th partial in order, then the fixed
// block tree.
__global__ __launch_bounds__(256) void k_residual_final(const double* __restrict__ part, int64_t nblk, int k,
                                                        double* out, int root) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int64_t q = threadIdx.x; q < nblk; q += 256) s += part[q * k + j];
    s = block_sum256(s, red);
    if (threadIdx.x == 0) out[j] = root ? sqrt(s) : s;
}

// B[j][i] = A[i][j] for an rows x cols block (32 x 32 tiles through LDS;
// grid x over the rows, which may be many, y over the columns).
__global__ __launch_bounds__(256) void k_transpose(int64_t rows, int64_t cols, const double* __restrict__ A,
                                                   int64_t lda, double* __restrict__ B, int64_t ldb) {
    __shared__ double t[32][33];
    const int64_t tr = (int64_t)blockIdx.x * 32, tc = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = tr + r, j = tc + tx;
        if (i < rows && j < cols) t[r][tx] = A[i * lda + j];
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int64_t j = tc + c, i = tr + tx;
        if (i < rows && j < cols) B[j * ldb + i] = t[tx][c];
    }
}

// The same for cols <= NB (a block of a few vectors): one thread per row
// reads its NB values and writes them to NB rows of B (coalesced over i).
template <int NB>
__global__ __launch_bounds__(256) void k_transpose_narrow(int64_t rows, int cols, const double* __restrict__ A,
                                                          int64_t lda, double* __restrict__ B, int64_t ldb) {
    XR_GRID_STRIDE(i, rows) {
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j < cols) B[j * ldb + i] = A[i * lda + j];
    }
}

static int residual_go(xr_ctx* c, int64_t n, int b, int k, const double* X, int64_t ldx, const double* A,
                       int64_t lda, const double* theta, double* R, int64_t ldr, double* out, int root) {
    const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256 * 4), 1024));
    double* part = nullptr;
    if (k > 0) {
        part = (double*)xr_workspace(c, WS_SOLVER_PART, (size_t)(nblk * k * sizeof(double)));
        if (!part) return xr_fail(c, XR_ENOMEM, "workspace (column norms)");
    }
    ProfScope prof(c, XR_K_OTHER, 0.0, 8.0 * n * b * (X ? 3 : 1));
#define XR_RES(NB_) k_residual_part<NB_><<<(unsigned)nblk, 256, 0, c->stream>>>(n, b, k, X, ldx, A, lda, theta, R, ldr, part)
    if (b <= 1) { XR_RES(1); }
    else if (b <= 2) { XR_RES(2); }
    else if (b <= 4) { XR_RES(4); }
    else if (b <= 8) { XR_RES(8); }
    else if (b <= 16) { XR_RES(16); }
    else { XR_RES(32); }
#undef XR_RES
    XR_LAUNCH_CHECK(c);
    if (k > 0) {
        k_residual_final<<<(unsigned)k, 256, 0, c->stream>>>(part, nblk, k, out, root);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

extern "C" int xr_precond_divide(xr_ctx* c, int64_t n, int64_t b, const double* X, int64_t ldx, double* Y,
                                 int64_t ldy, const double* lam, const double* theta, double sign, double floor_) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n < 0 || b < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (n == 0 || b == 0) return XR_OK;
    if (ldx < b || ldy < b) return xr_fail(c, XR_ESHAPE, "row stride below the block width");
    if (!(floor_ > 0.0)) return xr_fail(c, XR_EINVAL, "floor must be positive");
    ProfScope prof(c, XR_K_OTHER, 0.0, 16.0 * n * b);
    k_precond_divide<<<xr_grid(n * b), 256, 0, c->stream>>>(n, b, X, ldx, Y, ldy, lam, theta, sign, floor_);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

extern "C" int xr_residual(xr_ctx* c, int64_t n, int b, int k, const double* X, int64_t ldx, const double* AX,
                           int64_t ldax, const double* theta, double* R, int64_t ldr, double* norms) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n < 0 || b < 0 || k < 0 || k > b) return xr_fail(c, XR_EINVAL, "bad residual sizes (0 <= k <= b)");
    if (b > XR_SOLVER_MAXB) return xr_fail(c, XR_EUNSUPPORTED, "residual block wider than 32 columns");
    if (ldx < b || ldax < b || ldr < b) return xr_fail(c, XR_ESHAPE, "row stride below the block width");
    if (n == 0 || b == 0) return XR_OK;
    if (!X || !AX || !theta || !R || (k > 0 && !norms)) return xr_fail(c, XR_EINVAL, "null residual operand");
    return residual_go(c, n, b, k, X, ldx, AX, ldax, theta, R, ldr, norms, 1);
}

extern "C" int xr_column_norms(xr_ctx* c, int64_t n, int b, const double* A, int64_t lda, double* out,
                               int squared) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n < 0 || b < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (b > XR_SOLVER_MAXB) return xr_fail(c, XR_EUNSUPPORTED, "more than 32 columns");
    if (lda < b) return xr_fail(c, XR_ESHAPE, "row stride below the block width");
    if (b == 0) return XR_OK;
    if (!A || !out) return xr_fail(c, XR_EINVAL, "null operand");
    return residual_go(c, n, b, b, nullptr, 0, A, lda, nullptr, nullptr, 0, out, squared ? 0 : 1);
}

extern "C" int xr_transpose(xr_ctx* c, int64_t rows, int64_t cols, const double* A, int64_t lda, double* B,
                            int64_t ldb) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (rows < 0 || cols < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (rows == 0 || cols == 0) return XR_OK;
    if (lda < cols || ldb < rows) return xr_fail(c, XR_ESHAPE, "leading dimension too small");
    ProfScope prof(c, XR_K_OTHER, 0.0, 16.0 * rows * cols);
    if (cols <= 8) {
#define XR_TN(NB_) k_transpose_narrow<NB_><<<xr_grid(rows), 256, 0, c->stream>>>(rows, (int)cols, A, lda, B, ldb)
        if (cols <= 1) { XR_TN(1); }
        else if (cols <= 2) { XR_TN(2); }
        else if (cols <= 4) { XR_TN(4); }
        else { XR_TN(8); }
#undef XR_TN
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    if (ceil_div(rows, 32) > 0x7fffffff || ceil_div(cols, 32) > 65535)
        return xr_fail(c, XR_EUNSUPPORTED, "transpose grid too large");
    k_transpose<<<dim3((unsigned)ceil_div(rows, 32), (unsigned)ceil_div(cols, 32)), 256, 0, c->stream>>>(
        rows, cols, A, lda, B, ldb);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// ---------------------------------------------------------------------------
// Diagonal of the XR operator in the product eigenbasis, one sector tuple at
// a time: out[i] = sum_m e_m[j_m] + sum_(m<n) pd_mn[j_m, j_n] for the
// row-major digits (j_0, ..., j_{nf-1}) of i, the terms summed in that order.
// ---------------------------------------------------------------------------
#define XR_DIAG_MAXF 8
enum { WS_COLDOT = 93 };

struct DiagTuple {
    int nf;
    int64_t size;
    int64_t dims[XR_DIAG_MAXF];
    const double* e[XR_DIAG_MAXF];
    const double* pd[XR_DIAG_MAXF * (XR_DIAG_MAXF - 1) / 2];
};

__global__ __launch_bounds__(256) void k_sector_diag(DiagTuple t, double* __restrict__ out) {
    XR_GRID_STRIDE(i, t.size) {
        int64_t j[XR_DIAG_MAXF];
        int64_t r = i;
        for (int m = t.nf - 1; m >= 0; m--) {
            const int64_t q = r / t.dims[m];
            j[m] = r - q * t.dims[m];
            r = q;
        }
        double v = 0.0;
        for (int m = 0; m < t.nf; m++) v += t.e[m][j[m]];
        int p = 0;
        for (int m = 0; m < t.nf; m++)
            for (int n = m + 1; n < t.nf; n++, p++) v += t.pd[p][j[m] * t.dims[n] + j[n]];
        out[i] = v;
    }
}

extern "C" int xr_sector_diag(xr_ctx* c, int nf, int ntup, const int64_t* dims, const int64_t* off,
                              const double* const* e, const double* const* pd, double* out) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (nf < 1 || nf > XR_DIAG_MAXF) return xr_fail(c, XR_EUNSUPPORTED, "1 to 8 fragments");
    if (ntup < 0) return xr_fail(c, XR_EINVAL, "negative tuple count");
    if (ntup == 0) return XR_OK;
    if (!dims || !off || !e || (nf > 1 && !pd) || !out) return xr_fail(c, XR_EINVAL, "null argument");
    const int np = nf * (nf - 1) / 2;
    for (int q = 0; q < ntup; q++) {
        DiagTuple t{};
        t.nf = nf;
        t.size = 1;
        for (int m = 0; m < nf; m++) {
            const int64_t d = dims[(int64_t)q * nf + m];
            if (d < 1) return xr_fail(c, XR_EINVAL, "empty tuple axis");
            t.dims[m] = d;
            t.size *= d;
            t.e[m] = e[(int64_t)q * nf + m];
            if (!t.e[m]) return xr_fail(c, XR_EINVAL, "null monomer diagonal");
        }
        for (int p = 0; p < np; p++) {
            t.pd[p] = pd[(int64_t)q * np + p];
            if (!t.pd[p]) return xr_fail(c, XR_EINVAL, "null pair diagonal");
        }
        ProfScope prof(c, XR_K_OTHER, 0.0, 8.0 * t.size);
        k_sector_diag<<<xr_grid(t.size), 256, 0, c->stream>>>(t, out + off[q]);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

// out[j] = sum_i A[i, j] B[i, j] for rows x cols blocks: partial sums over
// chunks of 64 rows (one thread per column, in row order), then the chunks
// in order (deterministic).
__global__ __launch_bounds__(256) void k_coldot_part(int64_t rows, int64_t cols, const double* __restrict__ A,
                                                     int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                     double* __restrict__ part) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * 64;
    if (j >= cols) return;
    const int64_t r1 = r0 + 64 < rows ? r0 + 64 : rows;
    double s = 0.0;
    for (int64_t i = r0; i < r1; i++) s += A[i * lda + j] * B[i * ldb + j];
    part[(int64_t)blockIdx.y * cols + j] = s;
}

__global__ __launch_bounds__(256) void k_coldot_final(const double* __restrict__ part, int64_t nchunk, int64_t cols,
                                                      double* __restrict__ out) {
    XR_GRID_STRIDE(j, cols) {
        double s = 0.0;
        for (int64_t q = 0; q < nchunk; q++) s += part[q * cols + j];
        out[j] = s;
    }
}

extern "C" int xr_column_dots(xr_ctx* c, int64_t rows, int64_t cols, const double* A, int64_t lda, const double* B,
                              int64_t ldb, double* out) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (rows < 0 || cols < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (cols == 0) return XR_OK;
    if (lda < cols || ldb < cols) return xr_fail(c, XR_ESHAPE, "leading dimension below the column count");
    if (!A || !B || !out) return xr_fail(c, XR_EINVAL, "null operand");
    const int64_t nchunk = std::max<int64_t>(1, ceil_div(rows, 64));
    if (nchunk > 65535 || ceil_div(cols, 256) > 0x7fffffff) return xr_fail(c, XR_EUNSUPPORTED, "block too large");
    double* part = (double*)xr_workspace(c, WS_COLDOT, sizeof(double) * (size_t)(nchunk * cols));
    if (!part) return xr_fail(c, XR_ENOMEM, "workspace (column dots)");
    ProfScope prof(c, XR_K_OTHER, 2.0 * rows * cols, 16.0 * rows * cols);
    if (rows == 0) {
        k_coldot_final<<<xr_grid(cols), 256, 0, c->stream>>>(part, 0, cols, out);
    } else {
        k_coldot_part<<<dim3((unsigned)ceil_div(cols, 256), (unsigned)nchunk), 256, 0, c->stream>>>(rows, cols, A, lda,
                                                                                                   B, ldb, part);
        XR_LAUNCH_CHECK(c);
        k_coldot_final<<<xr_grid(cols), 256, 0, c->stream>>>(part, nchunk, cols, out);
    }
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}
