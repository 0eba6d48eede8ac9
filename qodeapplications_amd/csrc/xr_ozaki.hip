// Split-int8 FP64 GEMM on the int8 matrix cores (v_mfma_i32_16x16x64_i8).
//
// The Ozaki scheme: Ozaki, Ogita, Oishi, Rump, "Error-free transformations
// of matrix multiplication by using fast routines of matrix multiplication
// and its applications", Numer. Algorithms 59 (2012) 95-118; on integer
// matrix units: Ootomo, Ozaki, Yokota, "DGEMM on integer matrix
// multiplication unit", Int. J. High Perform. Comput. Appl. 38 (2024).
//
// Every row of op(A) and every column of op(B) (a "row" r of an operand
// below, with its own K values) gets the exponent E_r = the largest biased
// exponent field of its values, so |x| < 2^(E_r - 1022).  The scaled value
// y = x 2^(1028 - E_r) (|y| < 64) is cut into XR_I8_S signed digits:
//     d_1 = rint(y),  y <- 128 (y - d_i),  d_(i+1) = rint(y),  |d_i| <= 64,
//     x = 2^(E_r - 1022) (sum_i d_i 2^(-6 - 7 (i - 1)) + rho),  |rho| <= 2^(-7 S)
// (every step is exact in FP64).  A product then is
//     a b = 2^(E_a + E_b - 2056) sum_t 2^(-7 t) sum_(i + j = t) d^a_i d^b_j
// (0-based i, j) plus the rho terms.  The pairs with i + j <= S - 1 are
// formed on the int8 matrix cores; |d^a d^b| <= 2^12, so the int32 sum over
// one t and K values stays below K S 2^12 < 2^31 for K <= 65536 (longer K
// is split).  The epilogue is the only FP64 arithmetic: a Horner sum over t
// and the power-of-two scale.  The terms left out (pairs with i + j >= S and
// the rho terms) are at most (S + 2) 2^(-7 S) 2^(E_a + E_b - 2044) per k,
// so |C - AB| <= K (S + 2) 2^(-7 S) 4 max|a_r| max|b_c| elementwise; with
// S = 7 that is 36 K 2^-49 max|a_r| max|b_c|.  The products cost
// S (S + 1) / 2 = 28 int8 MFMAs per FP64-equivalent one, against an int8
// matrix rate 64x the FP64 one on this chip.
// ---------------------------------------------------------------------------
#define XR_I8_S 7
#define XR_I8_KMAX 65536
#define XR_I8_ECHUNK 512        // k values per partial row maximum
typedef int i4_t __attribute__((ext_vector_type(4)));

// One operand of a split-int8 product: R rows of K values (kc: a row's K
// values are contiguous, X[r ld + k]; else X[k ld + r]), batch stride sX.
// P: partial row maxima [batch][nke][Rp]; E: row exponents [batch][Rp];
// D: digit planes [batch][S][Rp][Kp] (int8).
struct I8Operand {
    const double* X;
    int64_t ld, sX, R, K, Rp, Kp, nke;
    int kc;
    int* P;
    int* E;
    int8_t* D;
};

// Partial row maxima of the biased exponent field over 64-row x
// XR_I8_ECHUNK-k tiles of both operands (block ids: operand A's tiles of
// every batch, then B's); every entry of P is written (0 for padding rows).
__global__ __launch_bounds__(256) void k_i8_rowmax(I8Operand a, I8Operand b, int64_t batch) {
    __shared__ int red[4][64];
    const int64_t nA = batch * (a.Rp / 64) * a.nke, total = nA + batch * (b.Rp / 64) * b.nke;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int64_t id = blockIdx.x; id < total; id += gridDim.x) {
        const bool isA = id < nA;
        const I8Operand& o = isA ? a : b;
        const int64_t j = isA ? id : id - nA;
        const int64_t per = (o.Rp / 64) * o.nke, bt = j / per, rem = j - bt * per;
        const int64_t rb = rem / o.nke, kb = rem - rb * o.nke;
        const double* X = o.X + bt * o.sX;
        const int64_t r0 = rb * 64, k0 = kb * XR_I8_ECHUNK;
        const int64_t kn = (o.K - k0 < XR_I8_ECHUNK) ? o.K - k0 : XR_I8_ECHUNK;
        int* P = o.P + (bt * o.nke + kb) * o.Rp + r0;
        if (o.kc) {
            for (int rr = 0; rr < 16; rr++) {
                const int64_t gr = r0 + w * 16 + rr;
                int e = 0;
                if (gr < o.R) {
                    const double* row = X + gr * o.ld + k0;
#pragma unroll
                    for (int q = 0; q < XR_I8_ECHUNK / 64; q++)
                        if (lane + 64 * q < kn) e = max(e, xr_expo(row[lane + 64 * q]));
                }
#pragma unroll
                for (int s = 32; s > 0; s >>= 1) e = max(e, __shfl_xor(e, s));
                if (lane == 0) P[w * 16 + rr] = e;
            }
        } else {
            const int64_t gr = r0 + lane;
            int e = 0;
            if (gr < o.R) {
                const double* col = X + k0 * o.ld + gr;
#pragma unroll 8
                for (int kk = w; kk < kn; kk += 4) e = max(e, xr_expo(col[kk * o.ld]));
            }
            red[w][lane] = e;
            __syncthreads();
            if (w == 0) P[lane] = max(max(red[0][lane], red[1][lane]), max(red[2][lane], red[3][lane]));
            __syncthreads();
        }
    }
}

// Row exponents E_r = max of the row's partials, for both operands.
__global__ __launch_bounds__(256) void k_i8_rowexp(I8Operand a, I8Operand b, int64_t batch) {
    const int64_t nA = batch * a.Rp;
    XR_GRID_STRIDE(i, nA + batch * b.Rp) {
        const bool isA = i < nA;
        const I8Operand& o = isA ? a : b;
        const int64_t j = isA ? i : i - nA, bt = j / o.Rp, r = j - bt * o.Rp;
        const int* P = o.P + bt * o.nke * o.Rp + r;
        int e = 0;
        for (int64_t q = 0; q < o.nke; q++) e = max(e, P[q * o.Rp]);
        o.E[j] = e;
    }
}

// Digit planes of 64 x 64 tiles of both operands' padded Rp x Kp layouts
// (zeros outside R x K), on the rows' exponents.
__global__ __launch_bounds__(256) void k_i8_slice(I8Operand a, I8Operand b, int64_t batch) {
    __shared__ double T[64][65];
    __shared__ int Es[64];
    const int64_t nA = batch * (a.Rp / 64) * (a.Kp / 64), total = nA + batch * (b.Rp / 64) * (b.Kp / 64);
    const int tid = threadIdx.x;
    for (int64_t id = blockIdx.x; id < total; id += gridDim.x) {
        const bool isA = id < nA;
        const I8Operand& o = isA ? a : b;
        const int64_t j = isA ? id : id - nA;
        const int64_t nkb = o.Kp / 64, per = (o.Rp / 64) * nkb, bt = j / per, rem = j - bt * per;
        const int64_t rb = rem / nkb, kb = rem - rb * nkb;
        const double* X = o.X + bt * o.sX;
        const int64_t r0 = rb * 64, k0 = kb * 64;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int e = tid + 256 * i, s = e >> 6, f = e & 63;
            const int r = o.kc ? s : f, k = o.kc ? f : s;
            const int64_t gr = r0 + r, gk = k0 + k;
            T[r][k] = (gr < o.R && gk < o.K) ? X[o.kc ? gr * o.ld + gk : gk * o.ld + gr] : 0.0;
        }
        if (tid < 64) Es[tid] = o.E[bt * o.Rp + r0 + tid];
        __syncthreads();
        const int r = tid >> 2, g = tid & 3;
        const int sh = 1028 - Es[r];
        int w[XR_I8_S][4];
#pragma unroll
        for (int i = 0; i < XR_I8_S; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) w[i][q] = 0;
#pragma unroll
        for (int jj = 0; jj < 16; jj++) {
            double y = ldexp(T[r][16 * g + jj], sh);
#pragma unroll
            for (int i = 0; i < XR_I8_S; i++) {
                const double d = rint(y);
                w[i][jj >> 2] |= ((int)d & 0xff) << (8 * (jj & 3));
                y = (y - d) * 128.0;
            }
        }
        int8_t* out = o.D + bt * XR_I8_S * o.Rp * o.Kp + (r0 + r) * o.Kp + k0 + 16 * g;
#pragma unroll
        for (int i = 0; i < XR_I8_S; i++)
            *reinterpret_cast<i4_t*>(out + (int64_t)i * o.Rp * o.Kp) = i4_t{w[i][0], w[i][1], w[i][2], w[i][3]};
        __syncthreads();
    }
}

// One 64 x 64 output tile over the k range [kbeg, kbeg + kchunk) of the
// digit planes.  4 waves in 2 x 2, each 32 x 32 as 2 x 2 MFMA tiles, with
// one int32 accumulator per digit-weight t.  Per 64-deep k step every digit
// plane of both operands is staged in LDS (64 rows x 64 bytes each, laid
// out so a fragment read of 64 lanes x 16 bytes is one contiguous 1 KiB
// run); each wave holds all B fragments and walks the A planes.  No
// register prefetch: the kernel fits two workgroups per CU (2 x 56 KiB of
// LDS, <= 256 registers a lane), and each one's loads overlap the other's
// matrix work.
template <int S>
__global__ __launch_bounds__(256, 2) void k_i8gemm(
    const int8_t* __restrict__ DA, const int* __restrict__ EA, int64_t RpA, const int8_t* __restrict__ DB,
    const int* __restrict__ EB, int64_t RpB, int64_t Kp, int64_t M, int64_t N, double alpha,
    double* __restrict__ C, int64_t ldc, int64_t sC, int64_t sSplit, int64_t tiles_m, int64_t tiles_n,
    int64_t kchunk) {
    __shared__ __align__(16) int8_t sA[S][64 * 64];
    __shared__ __align__(16) int8_t sB[S][64 * 64];
    const int64_t batch = blockIdx.z;
    DA += batch * S * RpA * Kp;
    DB += batch * S * RpB * Kp;
    EA += batch * RpA;
    EB += batch * RpB;
    C += batch * sC + (int64_t)blockIdx.y * sSplit;
    const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
    const int64_t kend = (Kp - kbeg < kchunk) ? Kp : kbeg + kchunk;
    const int64_t tile = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    const int64_t GROUP = 8, per_group = GROUP * tiles_n, gq = tile / per_group, first_m = gq * GROUP;
    const int64_t gsize = (tiles_m - first_m < GROUP) ? (tiles_m - first_m) : GROUP;
    const int64_t tm = first_m + (tile % per_group) % gsize, tn = (tile % per_group) / gsize;
    const int64_t m0 = tm * 64, n0 = tn * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
    const int lrow = tid >> 2, lg = tid & 3;
    const int so = ((lrow >> 4) * 4 + lg) * 256 + (lrow & 15) * 16;
    const int fa = ((wm * 2) * 4 + (lane >> 4)) * 256 + (lane & 15) * 16;
    const int fb = ((wn * 2) * 4 + (lane >> 4)) * 256 + (lane & 15) * 16;
    const int8_t* pa = DA + (m0 + lrow) * Kp + 16 * lg;
    const int8_t* pb = DB + (n0 + lrow) * Kp + 16 * lg;
    const int64_t planeA = RpA * Kp, planeB = RpB * Kp;
    auto stage = [&](int64_t k) {
        i4_t ga[S], gb[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            ga[i] = *reinterpret_cast<const i4_t*>(pa + i * planeA + k);
            gb[i] = *reinterpret_cast<const i4_t*>(pb + i * planeB + k);
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            *reinterpret_cast<i4_t*>(&sA[i][so]) = ga[i];
            *reinterpret_cast<i4_t*>(&sB[i][so]) = gb[i];
        }
    };
    i4_t acc[S][2][2];
#pragma unroll
    for (int t = 0; t < S; t++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[t][i][j] = i4_t{0, 0, 0, 0};
    const int64_t nk = (kend - kbeg) / 64;
    for (int64_t kt = 0; kt < nk; kt++) {
        stage(kbeg + kt * 64);
        __syncthreads();
        i4_t bf[S][2];
#pragma unroll
        for (int j = 0; j < S; j++) {
            bf[j][0] = *reinterpret_cast<const i4_t*>(&sB[j][fb]);
            bf[j][1] = *reinterpret_cast<const i4_t*>(&sB[j][fb + 1024]);
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            const i4_t a0 = *reinterpret_cast<const i4_t*>(&sA[i][fa]);
            const i4_t a1 = *reinterpret_cast<const i4_t*>(&sA[i][fa + 1024]);
#pragma unroll
            for (int j = 0; j + i < S; j++) {
                acc[i + j][0][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bf[j][0], acc[i + j][0][0], 0, 0, 0);
                acc[i + j][0][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bf[j][1], acc[i + j][0][1], 0, 0, 0);
                acc[i + j][1][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bf[j][0], acc[i + j][1][0], 0, 0, 0);
                acc[i + j][1][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bf[j][1], acc[i + j][1][1], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[row][col] of the 16 x 16 int32 MFMA: col = lane % 16, row = 4 (lane / 16) + r
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t gm = m0 + wm * 32 + bi * 16 + 4 * (lane >> 4) + r;
                const int64_t gn = n0 + wn * 32 + bj * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    double v = (double)acc[S - 1][bi][bj][r];
#pragma unroll
                    for (int t = S - 2; t >= 0; t--) v = v * 0.0078125 + (double)acc[t][bi][bj][r];
                    C[gm * ldc + gn] = alpha * ldexp(v, EA[gm] + EB[gn] - 2056);
                }
            }
}

enum { WS_I8_DA = 80, WS_I8_DB, WS_I8_E, WS_I8_C };

// Split-int8 product: digit planes of both operands (three launches), the
// int8 GEMM (K split toward a full chip and into chunks of at most
// XR_I8_KMAX), and the fixed-order reduction when the product is split or
// accumulates (beta).
int xr_dgemm_i8_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                       const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB,
                       double beta, double* C, int64_t ldc, int64_t sC, int64_t batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return XR_OK;
    if (K <= 0)
        return xr_dgemm_launch(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    const int64_t RpA = ceil_div(M, 64) * 64, RpB = ceil_div(N, 64) * 64, Kp = ceil_div(K, 64) * 64;
    const int64_t nke = ceil_div(K, XR_I8_ECHUNK);
    int64_t tm = ceil_div(M, 64), tn = ceil_div(N, 64), tiles = tm * tn;
    if (tiles > 0x7fffffff || batch > 65535) return xr_fail(c, XR_EUNSUPPORTED, "GEMM grid too large");
    int64_t splits = ceil_div(Kp, XR_I8_KMAX);
    if (tiles * batch < 512 && Kp >= 512) {
        int64_t s = ceil_div(1024, tiles * batch);
        s = std::min<int64_t>(std::min<int64_t>(s, Kp / 256), 512);
        splits = std::max(splits, s);
    }
    int64_t kchunk = ceil_div(ceil_div(Kp, splits), 64) * 64;
    splits = ceil_div(Kp, kchunk);
    int8_t* DA = (int8_t*)xr_workspace(c, WS_I8_DA, (size_t)(batch * XR_I8_S * RpA * Kp));
    int8_t* DB = (int8_t*)xr_workspace(c, WS_I8_DB, (size_t)(batch * XR_I8_S * RpB * Kp));
    int* E = (int*)xr_workspace(c, WS_I8_E, (size_t)(batch * (RpA + RpB) * (nke + 1) * sizeof(int)));
    const bool direct = (splits == 1 && beta == 0.0);
    double* Cw = direct ? C : (double*)xr_workspace(c, WS_I8_C, (size_t)(splits * batch * M * N * sizeof(double)));
    if (!DA || !DB || !E || !Cw) return xr_fail(c, XR_ENOMEM, "workspace (split-int8)");
    int* EA = E;
    int* EB = EA + batch * RpA;
    int* PA = EB + batch * RpB;
    int* PB = PA + batch * nke * RpA;
    const I8Operand oa{A, lda, sA, M, K, RpA, Kp, nke, !ta, PA, EA, DA};
    const I8Operand ob{B, ldb, sB, N, K, RpB, Kp, nke, tb != 0, PB, EB, DB};
    {
        ProfScope prof(c, XR_K_I8_SLICE, 0.0, 8.0 * (M * K + K * N) * batch * 2 + (double)XR_I8_S * (RpA + RpB) * Kp * batch);
        prof.shape(M, N, K, batch, ta, tb);
        const int64_t ne = batch * (RpA / 64 + RpB / 64) * nke, ns = batch * (RpA / 64 + RpB / 64) * (Kp / 64);
        k_i8_rowmax<<<(unsigned)std::min<int64_t>(ne, 1 << 20), 256, 0, c->stream>>>(oa, ob, batch);
        XR_LAUNCH_CHECK(c);
        k_i8_rowexp<<<xr_grid(batch * (RpA + RpB)), 256, 0, c->stream>>>(oa, ob, batch);
        XR_LAUNCH_CHECK(c);
        k_i8_slice<<<(unsigned)std::min<int64_t>(ns, 1 << 20), 256, 0, c->stream>>>(oa, ob, batch);
        XR_LAUNCH_CHECK(c);
    }
    const dim3 grid((unsigned)tiles, (unsigned)splits, (unsigned)batch);
    {
        ProfScope prof(c, XR_K_GEMM_SPLIT_I8, 2.0 * M * N * K * batch,
                       (double)XR_I8_S * (RpA + RpB) * Kp * batch + 8.0 * M * N * batch * splits);
        prof.shape(M, N, K, batch, ta, tb);
        if (direct)
            k_i8gemm<XR_I8_S><<<grid, 256, 0, c->stream>>>(DA, EA, RpA, DB, EB, RpB, Kp, M, N, alpha, C, ldc, sC, 0,
                                                          tm, tn, kchunk);
        else
            k_i8gemm<XR_I8_S><<<grid, 256, 0, c->stream>>>(DA, EA, RpA, DB, EB, RpB, Kp, M, N, 1.0, Cw, N, M * N,
                                                          batch * M * N, tm, tn, kchunk);
        XR_LAUNCH_CHECK(c);
    }
    if (direct) return XR_OK;
    return splitk_reduce(c, Cw, splits, batch, M, N, alpha, beta, C, ldc, sC);
}

// Whether a product of this shape gains from the split-int8 path (measured
// on the XR shapes): at least 2^33 flops, so the three digit-plane launches
// and passes stay small beside it; every dimension at least 128 (narrower
// ones leave the 64-wide tiles and the digit passes mostly idle); and M or
// N at least 1024 (a small output over a long K is bound by re-reading the
// 7 digit planes through L2, where the FP64 kernel's larger tiles win).
// Other products of the split-int8 mode run on the FP64 matrix cores.
bool xr_i8_pays(int64_t M, int64_t N, int64_t K, int64_t batch) {
    if (M < 128 || N < 128 || K < 128) return false;
    return (M >= 1024 || N >= 1024) && 2.0 * M * N * K * batch >= 8589934592.0;
}

extern "C" int xr_gemm_split_i8(xr_ctx* c, int transA, int transB, int64_t M, int64_t N, int64_t K,
                                double alpha, const double* A, int64_t lda, int64_t sA, const double* B,
                                int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc, int64_t sC,
                                int64_t batch) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M < 0 || N < 0 || K < 0 || batch < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
    if (ldc < N) return xr_fail(c, XR_ESHAPE, "ldc < N");
    if ((transA ? lda < M : lda < K) && M > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "lda too small");
    if ((transB ? ldb < K : ldb < N) && N > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "ldb too small");
    return xr_dgemm_i8_launch(c, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

extern "C" int xr_gemm_mode(xr_ctx* c, int mode, int transA, int transB, int64_t M, int64_t N, int64_t K,
                            double alpha, const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                            int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M < 0 || N < 0 || K < 0 || batch < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
    if (ldc < N) return xr_fail(c, XR_ESHAPE, "ldc < N");
    if ((transA ? lda < M : lda < K) && M > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "lda too small");
    if ((transB ? ldb < K : ldb < N) && N > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "ldb too small");
    if (mode != XR_PREC_F64 && mode != XR_PREC_SPLIT_F32 && mode != XR_PREC_SPLIT_I8)
        return xr_fail(c, XR_EINVAL, "unknown precision mode");
    const int saved = c->precision;
    c->precision = mode;
    const int rc = xr_gemm_2e(c, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    c->precision = saved;
    return rc;
}
