// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C_b = alpha * op(A_b) op(B_b) + beta * C_b.  The tile kernel computes
// alpha * op(A) op(B) only (alpha is applied while A is staged in LDS) and
// stores the accumulators untouched: any arithmetic on them in the epilogue
// makes the compiler copy all 128 accumulator registers out of the AGPRs,
// which halves occupancy.  Split-K goes through a workspace and a
// fixed-order reduction, and so does beta != 0 unless the accumulating
// variant (k_dgemm_acc) costs nothing extra: small grids or short K.  256 threads = 4 waves per
// workgroup in a 2 x 2 arrangement; each wave owns a (BM/2) x (BN/2) block of
// 16 x 16 MFMA tiles.  K advances 16 at a time through one LDS stage while the
// next stage is prefetched into registers.  Workgroups are remapped so each
// XCD (own L2) receives a contiguous run of output tiles.
#include "xr_common.h"
#include "xr_gemm.h"
#include <vector>
#include <algorithm>

typedef double d4_t __attribute__((ext_vector_type(4)));

#define GEMM_BK 16
#define GEMM_THREADS 256
// output row of accumulator element r for lane group fk (lane / 16) of
// v_mfma_f32_16x16x4f32 (checked on gfx950; the f64 variant differs)
#define XR_F32_ROW(fk, r) (4 * (fk) + (r))

// blockIdx.x -> logical tile id, so that the ids one XCD executes (hardware
// hands out workgroups round-robin over the 8 XCDs) are contiguous.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t total) {
    const int64_t nx = 8;
    int64_t q = total / nx, r = total % nx;
    int64_t x = b % nx, j = b / nx;
    int64_t base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return base + j;
}

// Staged operand tiles are k-major with rows of BM (BN) + 16 doubles, so
// rows k and k + 1 fall in opposite halves of the 64 LDS banks and the MFMA
// fragment reads (16 lanes over the column, rows k and k + 1 per half-wave)
// touch 32 distinct 8-byte bank pairs.  An operand that arrives K-fastest
// from global memory (A not transposed, B transposed) is stored with 16
// lanes over k: there the column's bits 1..3 are XORed with (k >> 1) & 7,
// which spreads those stores over distinct bank pairs too; the reads apply
// the same XOR (its ks part is a compile-time constant of the unrolled loop).
__device__ __forceinline__ int lds_swz(int k, int col) {
    return col ^ (((k >> 1) & 7) << 1);
}

// One BM x BN output tile at (m0, n0): C = alpha op(A) op(B) over [0, K),
// or C = alpha op(A) op(B) + beta C with ACC (an epilogue that reads C costs
// the accumulators' AGPR residency, so only the accumulating variant has it).
template <int BM, int BN, bool TA, bool TB, bool ACC = false>
__device__ __forceinline__ void dgemm_tile(int64_t M, int64_t N, int64_t K, double alpha,
                                           const double* __restrict__ A, int64_t lda,
                                           const double* __restrict__ B, int64_t ldb,
                                           double* __restrict__ C, int64_t ldc, int64_t m0, int64_t n0,
                                           double beta = 0.0) {
    constexpr int WM = BM / 2, WN = BN / 2;    // wave tile
    constexpr int FM = WM / 16, FN = WN / 16;  // MFMA tiles per wave
    constexpr int A_PER = BM * GEMM_BK / GEMM_THREADS;
    constexpr int B_PER = BN * GEMM_BK / GEMM_THREADS;
    constexpr int SA = BM + 16, SB = BN + 16;  // LDS row strides (doubles)
    constexpr bool XA = !TA, XB = TB;          // operands stored with the column XOR
    __shared__ double As[GEMM_BK * SA];
    __shared__ double Bs[GEMM_BK * SB];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    double ra[A_PER], rb[B_PER];

    auto load_a = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < A_PER; i++) {
            int mm, kk;
            if (TA) {  // A[k][m]
                mm = tid % BM;
                kk = tid / BM + i * (GEMM_THREADS / BM);
            } else {   // A[m][k]
                kk = tid % GEMM_BK;
                mm = tid / GEMM_BK + i * (GEMM_THREADS / GEMM_BK);
            }
            int64_t gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < M && gk < K) v = TA ? A[gk * lda + gm] : A[gm * lda + gk];
            ra[i] = v;
        }
    };
    auto load_b = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < B_PER; i++) {
            int nn, kk;
            if (!TB) {  // B[k][n]
                nn = tid % BN;
                kk = tid / BN + i * (GEMM_THREADS / BN);
            } else {    // B[n][k]
                kk = tid % GEMM_BK;
                nn = tid / GEMM_BK + i * (GEMM_THREADS / GEMM_BK);
            }
            int64_t gn = n0 + nn, gk = k0 + kk;
            double v = 0.0;
            if (gn < N && gk < K) v = TB ? B[gn * ldb + gk] : B[gk * ldb + gn];
            rb[i] = v;
        }
    };
    auto store_ab = [&]() {
#pragma unroll
        for (int i = 0; i < A_PER; i++) {
            int mm, kk;
            if (TA) { mm = tid % BM; kk = tid / BM + i * (GEMM_THREADS / BM); }
            else { kk = tid % GEMM_BK; mm = tid / GEMM_BK + i * (GEMM_THREADS / GEMM_BK); }
            As[kk * SA + (XA ? lds_swz(kk, mm) : mm)] = alpha * ra[i];
        }
#pragma unroll
        for (int i = 0; i < B_PER; i++) {
            int nn, kk;
            if (!TB) { nn = tid % BN; kk = tid / BN + i * (GEMM_THREADS / BN); }
            else { kk = tid % GEMM_BK; nn = tid / GEMM_BK + i * (GEMM_THREADS / GEMM_BK); }
            Bs[kk * SB + (XB ? lds_swz(kk, nn) : nn)] = rb[i];
        }
    };

    d4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

    const int64_t nk = (K + GEMM_BK - 1) / GEMM_BK;
    if (nk > 0) {
        load_a(0);
        load_b(0);
        store_ab();
        __syncthreads();
    }
    const int fr = lane & 15;   // row / column inside a 16-wide fragment
    const int fk = lane >> 4;   // k inside a 4-deep step
    // fragment read offsets per 4-deep step ks: row ks*4 + fk, column
    // (w*W + i*16 + fr) with the XOR of row (ks*4 + fk), whose ks part is
    // 4 ks on the column's bits 2..3
    int aoff[GEMM_BK / 4], boff[GEMM_BK / 4];
#pragma unroll
    for (int ks = 0; ks < GEMM_BK / 4; ks++) {
        const int x = ((2 * ks + (fk >> 1)) & 7) << 1;
        aoff[ks] = fk * SA + wm * WM + (XA ? (fr ^ x) : fr);
        boff[ks] = fk * SB + wn * WN + (XB ? (fr ^ x) : fr);
    }
    for (int64_t kt = 0; kt < nk; kt++) {
        if (kt + 1 < nk) {
            load_a((kt + 1) * GEMM_BK);
            load_b((kt + 1) * GEMM_BK);
        }
#pragma unroll
        for (int ks = 0; ks < GEMM_BK / 4; ks++) {
            double af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; i++) af[i] = As[ks * 4 * SA + aoff[ks] + i * 16];
#pragma unroll
            for (int j = 0; j < FN; j++) bf[j] = Bs[ks * 4 * SB + boff[ks] + j * 16];
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (kt + 1 < nk) {
            store_ab();
            __syncthreads();
        }
    }

    // D[row][col] of a 16x16 tile (v_mfma_f64_16x16x4f64, checked on gfx950):
    // col = lane % 16, row = lane / 16 + 4 * r for accumulator element r.
    // With ACC the C values of one row of MFMA tiles are all loaded before
    // any is stored: loads and stores through the one C pointer would
    // otherwise stay in program order, one memory round trip per element.
#pragma unroll
    for (int i = 0; i < FM; i++) {
        double cv[FN][4];
        if (ACC) {
#pragma unroll
            for (int j = 0; j < FN; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t gm = m0 + wm * WM + i * 16 + fk + 4 * r;
                    const int64_t gn = n0 + wn * WN + j * 16 + fr;
                    cv[j][r] = (gm < M && gn < N) ? C[gm * ldc + gn] : 0.0;
                }
        }
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t gm = m0 + wm * WM + i * 16 + fk + 4 * r;
                const int64_t gn = n0 + wn * WN + j * 16 + fr;
                if (gm < M && gn < N) {
                    if (ACC) C[gm * ldc + gn] = acc[i][j][r] + beta * cv[j][r];
                    else C[gm * ldc + gn] = acc[i][j][r];
                }
            }
    }
}


template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(GEMM_THREADS) void k_dgemm(
    int64_t M, int64_t N, int64_t K, double alpha, const double* __restrict__ A, int64_t lda,
    int64_t sA, const double* __restrict__ B, int64_t ldb, int64_t sB, double beta,
    double* __restrict__ C, int64_t ldc, int64_t sC, int64_t tiles_m, int64_t tiles_n,
    int64_t kchunk, int64_t sSplit, int upper = 0) {
    (void)beta;
    const int64_t batch = blockIdx.z;
    A += batch * sA;
    B += batch * sB;
    C += batch * sC + (int64_t)blockIdx.y * sSplit;
    // split K: this workgroup covers [kbeg, kbeg + K) of the full K range
    const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
    K = (K - kbeg < kchunk) ? (K - kbeg) : kchunk;
    if (TA) A += kbeg * lda; else A += kbeg;
    if (TB) B += kbeg; else B += kbeg * ldb;
    const int64_t tile = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    // groups of 8 tile rows walk the tile columns together (L2 reuse of B)
    const int64_t GROUP = 8;
    const int64_t per_group = GROUP * tiles_n;
    const int64_t g = tile / per_group;
    const int64_t first_m = g * GROUP;
    const int64_t gsize = (tiles_m - first_m < GROUP) ? (tiles_m - first_m) : GROUP;
    const int64_t tm = first_m + (tile % per_group) % gsize;
    const int64_t tn = (tile % per_group) / gsize;
    if (upper && tm > tn) return;              // symmetric product: the upper tiles only
    const int64_t m0 = tm * BM, n0 = tn * BN;

    dgemm_tile<BM, BN, TA, TB>(M, N, K, alpha, A, lda, B, ldb, C, ldc, m0, n0);
}

// C_b = alpha op(A_b) op(B_b) + beta C_b without a K split, accumulating in
// the epilogue (saves the workspace pass of the staged path for beta != 0).
template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(GEMM_THREADS) void k_dgemm_acc(
    int64_t M, int64_t N, int64_t K, double alpha, const double* __restrict__ A, int64_t lda,
    int64_t sA, const double* __restrict__ B, int64_t ldb, int64_t sB, double beta,
    double* __restrict__ C, int64_t ldc, int64_t sC, int64_t tiles_m, int64_t tiles_n) {
    const int64_t batch = blockIdx.z;
    A += batch * sA;
    B += batch * sB;
    C += batch * sC;
    const int64_t tile = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    const int64_t GROUP = 8;
    const int64_t per_group = GROUP * tiles_n;
    const int64_t g = tile / per_group;
    const int64_t first_m = g * GROUP;
    const int64_t gsize = (tiles_m - first_m < GROUP) ? (tiles_m - first_m) : GROUP;
    const int64_t tm = first_m + (tile % per_group) % gsize;
    const int64_t tn = (tile % per_group) / gsize;
    dgemm_tile<BM, BN, TA, TB, true>(M, N, K, alpha, A, lda, B, ldb, C, ldc, tm * BM, tn * BN, beta);
}

// Grouped GEMM: up to XR_GROUP_PROBS independent products C_q = alpha
// op(A_q) op(B_q) of uneven shapes in one launch, so the launch fills the
// chip evenly where the problems alone would leave a partly empty last
// round of workgroups.  Workgroups are handed to the 8 XCDs round-robin
// (blockIdx % 8); every problem's tiles are split evenly over the XCDs, each
// XCD taking a contiguous run of them (L2 reuse inside the run), and each
// XCD works through the problems in order: per-tile work differs between
// problems (K), and this keeps the XCDs' loads equal.  The host orders the
// problems by decreasing K, so the longest tiles start first.
#define XR_GROUP_PROBS 32
struct GemmProb {
    const double* A;
    const double* B;
    double* C;
    int64_t M, N, K, lda, ldb, ldc, tiles_m, tiles_n;
};
struct GemmGroup {
    int n;
    int lower;           // square products whose strictly upper tiles are skipped
    double alpha, beta;  // beta: the accumulating variant only
    GemmProb p[XR_GROUP_PROBS];
};

template <int BM, int BN, bool TA, bool TB, bool ACC = false>
__global__ __launch_bounds__(GEMM_THREADS) void k_dgemm_grouped(GemmGroup g) {
    const int x = (int)(blockIdx.x & 7);
    int64_t j = blockIdx.x >> 3, local = -1;
    int q = 0;
    for (; q < g.n; q++) {                       // this XCD's share of each problem, in order
        const int64_t T = g.lower ? g.p[q].tiles_m * (g.p[q].tiles_m + 1) / 2 : g.p[q].tiles_m * g.p[q].tiles_n;
        const int64_t base = T >> 3, r = T & 7;
        const int64_t share = base + (x < r ? 1 : 0);
        if (j < share) {
            local = x * base + (x < r ? x : r) + j;
            break;
        }
        j -= share;
    }
    if (local < 0) return;                       // padding workgroup (uneven XCD shares)
    const GemmProb& P = g.p[q];
    int64_t tm, tn;
    if (g.lower) {                               // the lower-triangle tiles (tm >= tn) row by row
        tm = (int64_t)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
        while ((tm + 1) * (tm + 2) / 2 <= local) tm++;
        while (tm * (tm + 1) / 2 > local) tm--;
        tn = local - tm * (tm + 1) / 2;
    } else {
        const int64_t GROUP = 8;
        const int64_t per_group = GROUP * P.tiles_n;
        const int64_t gi = local / per_group;
        const int64_t first_m = gi * GROUP;
        const int64_t gsize = (P.tiles_m - first_m < GROUP) ? (P.tiles_m - first_m) : GROUP;
        tm = first_m + (local % per_group) % gsize;
        tn = (local % per_group) / gsize;
    }
    dgemm_tile<BM, BN, TA, TB, ACC>(P.M, P.N, P.K, g.alpha, P.A, P.lda, P.B, P.ldb, P.C, P.ldc, tm * BM, tn * BN,
                                    g.beta);
}

// C = alpha * sum_s P[s * step] + beta * C, summed in split order.
__global__ void k_splitk_reduce(const double* __restrict__ P, int64_t splits, int64_t batch,
                                int64_t M, int64_t N, double alpha, double beta,
                                double* __restrict__ C, int64_t ldc, int64_t sC, int64_t step = 1) {
    const int64_t n = batch * M * N;
    XR_GRID_STRIDE(i, n) {
        int64_t b = i / (M * N), r = i - b * M * N;
        int64_t m = r / N, col = r - m * N;
        double acc = 0.0;
        for (int64_t s = 0; s < splits; s += step) acc += P[s * n + i];
        double* cp = C + b * sC + m * ldc + col;
        double v = alpha * acc;
        if (beta != 0.0) v += beta * (*cp);
        *cp = v;
    }
}

// First stage of a reduction over many splits on a small output: each
// group of SPLIT_FOLD consecutive partials is summed (in order) into its
// first one, so the final pass reads every SPLIT_FOLD-th partial.
#define SPLIT_FOLD 32
__global__ void k_splitk_fold(double* __restrict__ P, int64_t splits, int64_t n) {
    const int64_t groups = (splits + SPLIT_FOLD - 1) / SPLIT_FOLD;
    XR_GRID_STRIDE(i, groups * n) {
        const int64_t g = i / n, e = i - g * n;
        const int64_t s0 = g * SPLIT_FOLD, s1 = s0 + SPLIT_FOLD < splits ? s0 + SPLIT_FOLD : splits;
        double acc = 0.0;
        for (int64_t s = s0; s < s1; s++) acc += P[s * n + e];
        P[s0 * n + e] = acc;
    }
}

// C = alpha * (sum of the split partials in Cw) + beta * C: outputs that
// leave most of the chip idle fold groups of partials first.
static int splitk_reduce(xr_ctx* c, double* Cw, int64_t splits, int64_t batch, int64_t M, int64_t N,
                         double alpha, double beta, double* C, int64_t ldc, int64_t sC) {
    const int64_t n = batch * M * N;
    int64_t step = 1;
    if (splits > 2 * SPLIT_FOLD && n < (1 << 18)) {
        k_splitk_fold<<<xr_grid(ceil_div(splits, SPLIT_FOLD) * n), 256, 0, c->stream>>>(Cw, splits, n);
        XR_LAUNCH_CHECK(c);
        step = SPLIT_FOLD;
    }
    k_splitk_reduce<<<xr_grid(n), 256, 0, c->stream>>>(Cw, splits, batch, M, N, alpha, beta, C, ldc, sC, step);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Tall-skinny product for N <= 8 with A row-major (M x K, the Davidson
// H X): memory bound on A, so each wave streams 4 rows of A once, reusing
// every loaded row of B for all 4, and reduces across lanes at the end.
template <int NB, bool TB>
__global__ __launch_bounds__(256) void k_dgemm_skinny(int64_t M, int64_t N, int64_t K, double alpha,
                                                      const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb,
                                                      double beta, double* __restrict__ C,
                                                      int64_t ldc) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r0 = wave * 4; r0 < M; r0 += nwaves * 4) {
        double acc[4][NB];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[r][j] = 0.0;
        const int nr = (M - r0 < 4) ? (int)(M - r0) : 4;
        for (int64_t k = lane; k < K; k += 64) {
            double b[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) b[j] = (j < N) ? (TB ? B[j * ldb + k] : B[k * ldb + j]) : 0.0;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double a = (r < nr) ? A[(r0 + r) * lda + k] : 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) acc[r][j] = fma(a, b[j], acc[r][j]);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double v = acc[r][j];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                acc[r][j] = v;
            }
        if (lane == 0) {
            for (int r = 0; r < nr; r++)
                for (int j = 0; j < N; j++) {
                    double* cp = C + (r0 + r) * ldc + j;
                    double v = alpha * acc[r][j];
                    if (beta != 0.0) v += beta * (*cp);
                    *cp = v;
                }
        }
    }
}

// C = alpha A^T B + beta C for A (K x M) and B (K x N) row-major with K
// large and M <= 64, N <= 8 (Gram products of a tall block of vectors).
// Each block sums a contiguous run of rows (lane = column of A, the 4 waves
// take every 4th row) into a partial (64 x NB) tile; a second kernel adds
// the partials in block order (deterministic).
template <int NB>
__global__ __launch_bounds__(256) void k_dgemm_tn_skinny(int64_t M, int64_t N, int64_t K,
                                                         const double* __restrict__ A, int64_t lda,
                                                         const double* __restrict__ B, int64_t ldb,
                                                         int64_t rows, double* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t k0 = (int64_t)blockIdx.x * rows;
    const int64_t k1 = (K - k0 < rows) ? K : k0 + rows;
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) acc[j] = 0.0;
#pragma unroll 4
    for (int64_t k = k0 + wave; k < k1; k += 4) {
        const double a = (lane < M) ? A[k * lda + lane] : 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) acc[j] = fma(a, (j < N) ? B[k * ldb + j] : 0.0, acc[j]);
    }
    __shared__ double red[4][64][NB];
#pragma unroll
    for (int j = 0; j < NB; j++) red[wave][lane][j] = acc[j];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int j = 0; j < NB; j++)
            part[((int64_t)blockIdx.x * 64 + lane) * NB + j] =
                red[0][lane][j] + red[1][lane][j] + red[2][lane][j] + red[3][lane][j];
    }
}

// C[i, j] = alpha * sum_b part[b][i][j] + beta C: one workgroup per output
// entry, the threads stride over the partial tiles and a fixed butterfly
// adds their sums (same order on every run).
__global__ __launch_bounds__(256) void k_tn_reduce(const double* __restrict__ part, int64_t nblk, int64_t M,
                                                   int64_t N, int nb, double alpha, double beta,
                                                   double* __restrict__ C, int64_t ldc) {
    const int64_t t = blockIdx.x;
    const int64_t i = t / N, j = t - (t / N) * N;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) s += part[(b * 64 + i) * nb + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = alpha * ((red[0] + red[1]) + (red[2] + red[3]));
        if (beta != 0.0) v += beta * C[i * ldc + j];
        C[i * ldc + j] = v;
    }
}

// Gram products with M <= 8 columns: the rows are spread over all 256
// threads (coalesced reads of A and B), each thread keeps the M x NB sums,
// and the block adds them through LDS into the same partial layout as above.
template <int MB, int NB>
__global__ __launch_bounds__(256) void k_dgemm_tn_small(int64_t M, int64_t N, int64_t K,
                                                        const double* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ B, int64_t ldb,
                                                        int64_t rows, double* __restrict__ part) {
    const int64_t k0 = (int64_t)blockIdx.x * rows;
    const int64_t k1 = (K - k0 < rows) ? K : k0 + rows;
    double acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = 0.0;
#pragma unroll 4
    for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
        double a[MB], b[NB];
#pragma unroll
        for (int i = 0; i < MB; i++) a[i] = (i < M) ? A[k * lda + i] : 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) b[j] = (j < N) ? B[k * ldb + j] : 0.0;
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __shared__ double red[256];
    for (int i = 0; i < MB; i++)
        for (int j = 0; j < NB; j++) {
            double v = acc[i][j];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
            __syncthreads();
            if (threadIdx.x == 0)
                part[((int64_t)blockIdx.x * 64 + i) * NB + j] = red[0] + red[1] + red[2] + red[3];
            __syncthreads();
        }
}

// C = alpha A op(B) + beta C for tall A (M x K row-major, row stride lda)
// with K <= 64 and N <= 8.  A block stages 64 rows of A through LDS with
// coalesced loads (the rows' K values are contiguous), then 4 threads per
// row sum interleaved k and combine with two shuffles.
template <int NB, bool TB>
__global__ __launch_bounds__(256) void k_dgemm_rows_lds(int64_t M, int64_t N, int64_t K, double alpha,
                                                        const double* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ B, int64_t ldb,
                                                        double beta, double* __restrict__ C, int64_t ldc) {
    __shared__ double tile[64 * 64];
    __shared__ double bs[64 * NB];
    const int tid = threadIdx.x;
    const int Ki = (int)K;
    for (int e = tid; e < Ki * NB; e += 256) {
        const int k = e / NB, j = e - (e / NB) * NB;
        bs[e] = (j < N) ? (TB ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j]) : 0.0;
    }
    const int row = tid >> 2, part = tid & 3;
    for (int64_t r0 = (int64_t)blockIdx.x * 64; r0 < M; r0 += (int64_t)gridDim.x * 64) {
        const int nrows = (M - r0 < 64) ? (int)(M - r0) : 64;
        __syncthreads();
        for (int e = tid; e < nrows * Ki; e += 256) {
            const int r = e / Ki, k = e - (e / Ki) * Ki;
            tile[e] = A[(r0 + r) * lda + k];
        }
        __syncthreads();
        double acc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) acc[j] = 0.0;
        if (row < nrows) {
            for (int k = part; k < Ki; k += 4) {
                const double a = tile[row * Ki + k];
#pragma unroll
                for (int j = 0; j < NB; j++) acc[j] = fma(a, bs[k * NB + j], acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            acc[j] += __shfl_xor(acc[j], 1, 64);
            acc[j] += __shfl_xor(acc[j], 2, 64);
        }
        if (part == 0 && row < nrows) {
#pragma unroll
            for (int j = 0; j < NB; j++) {
                if (j < N) {
                    double v = alpha * acc[j];
                    if (beta != 0.0) v += beta * C[(r0 + row) * ldc + j];
                    C[(r0 + row) * ldc + j] = v;
                }
            }
        }
    }
}

// C = alpha A op(B) + beta C for tall A (M x K row-major) with K <= 4 and
// N <= 8: one thread per output row (the K values of consecutive rows are
// adjacent, so a wave reads one contiguous run).
template <int NB, bool TB>
__global__ __launch_bounds__(256) void k_dgemm_rows_small(int64_t M, int64_t N, int64_t K, double alpha,
                                                          const double* __restrict__ A, int64_t lda,
                                                          const double* __restrict__ B, int64_t ldb,
                                                          double beta, double* __restrict__ C, int64_t ldc) {
    double bs[4][NB];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < NB; j++)
            bs[k][j] = (k < K && j < N) ? (TB ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j]) : 0.0;
    XR_GRID_STRIDE(r, M) {
        double acc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) acc[j] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < K) {
                const double a = A[r * lda + k];
#pragma unroll
                for (int j = 0; j < NB; j++) acc[j] = fma(a, bs[k][j], acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < N) {
                double v = alpha * acc[j];
                if (beta != 0.0) v += beta * C[r * ldc + j];
                C[r * ldc + j] = v;
            }
        }
    }
}

// C = alpha A^T op(B) + beta C for A stored K x M (row-major, lda) with
// K <= 64 and M large, N <= 8: the product of a block of vectors stored one
// per row with a small coefficient matrix.  One thread per output row reads
// column r of A, so consecutive threads read consecutive addresses.
template <int NB, bool TB>
__global__ __launch_bounds__(256) void k_dgemm_tn_tall(int64_t M, int64_t N, int64_t K, double alpha,
                                                       const double* __restrict__ A, int64_t lda,
                                                       const double* __restrict__ B, int64_t ldb,
                                                       double beta, double* __restrict__ C, int64_t ldc) {
    __shared__ double bs[64 * NB];
    for (int e = threadIdx.x; e < (int)K * NB; e += 256) {
        const int k = e / NB, j = e - (e / NB) * NB;
        bs[e] = (j < N) ? (TB ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j]) : 0.0;
    }
    __syncthreads();
    XR_GRID_STRIDE(r, M) {
        double acc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) acc[j] = 0.0;
        for (int k = 0; k < (int)K; k++) {
            const double a = A[(int64_t)k * lda + r];
#pragma unroll
            for (int j = 0; j < NB; j++) acc[j] = fma(a, bs[k * NB + j], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < N) {
                double v = alpha * acc[j];
                if (beta != 0.0) v += beta * C[r * ldc + j];
                C[r * ldc + j] = v;
            }
        }
    }
}

// C = alpha A B + beta C for A (M x K row-major) with M <= 64 / NB, N <= NB
// and K large (Gram products of vectors stored one per row).  Each block
// sums a contiguous run of k; wave w owns rows [w RB, w RB + RB) and its
// lanes read consecutive k of those rows (few accumulators per thread, so
// many waves stay resident to keep loads in flight).  The per-block sums go
// to a partial tile; k_tn_reduce adds the partials in block order.
template <int RB, int NB>
__global__ __launch_bounds__(256) void k_dgemm_nn_gram(int64_t M, int64_t N, int64_t K,
                                                       const double* __restrict__ A, int64_t lda,
                                                       const double* __restrict__ B, int64_t ldb,
                                                       int64_t rows, double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int i0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * RB;
    const int64_t k0 = (int64_t)blockIdx.x * rows;
    const int64_t k1 = (K - k0 < rows) ? K : k0 + rows;
    double acc[RB][NB];
#pragma unroll
    for (int r = 0; r < RB; r++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[r][j] = 0.0;
#pragma unroll 2
    for (int64_t k = k0 + lane; k < k1; k += 64) {
        double b[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) b[j] = (j < N) ? B[k * ldb + j] : 0.0;
#pragma unroll
        for (int r = 0; r < RB; r++) {
            if (i0 + r < M) {
                const double a = A[(int64_t)(i0 + r) * lda + k];
#pragma unroll
                for (int j = 0; j < NB; j++) acc[r][j] = fma(a, b[j], acc[r][j]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RB; r++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double v = acc[r][j];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0 && i0 + r < M) part[((int64_t)blockIdx.x * 64 + i0 + r) * NB + j] = v;
        }
}

static int launch_tn_skinny(xr_ctx* c, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                            int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                            int64_t ldc) {
    const int nb = N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : 8));
    int64_t nblk = ceil_div(K, 2048);
    if (nblk > 1024) nblk = 1024;
    const int64_t rows = ceil_div(K, nblk);
    nblk = ceil_div(K, rows);
    double* part = (double*)xr_workspace(c, 52, nblk * 64 * nb * sizeof(double));
    if (!part) return xr_fail(c, XR_ENOMEM, "workspace (Gram)");
    ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
    prof.shape(M, N, K, 1, 1, 0);
#define XR_TN_SMALL(MB_)                                                                                  \
    if (nb == 1) k_dgemm_tn_small<MB_, 1><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part); \
    else if (nb == 2) k_dgemm_tn_small<MB_, 2><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part); \
    else if (nb == 4) k_dgemm_tn_small<MB_, 4><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part); \
    else k_dgemm_tn_small<MB_, 8><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part)
    if (M <= 1) { XR_TN_SMALL(1); }
    else if (M <= 2) { XR_TN_SMALL(2); }
    else if (M <= 4) { XR_TN_SMALL(4); }
    else if (M <= 8) { XR_TN_SMALL(8); }
    else if (nb == 1) k_dgemm_tn_skinny<1><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part);
    else if (nb == 2) k_dgemm_tn_skinny<2><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part);
    else if (nb == 4) k_dgemm_tn_skinny<4><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part);
    else k_dgemm_tn_skinny<8><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part);
#undef XR_TN_SMALL
    XR_LAUNCH_CHECK(c);
    k_tn_reduce<<<(unsigned)(M * N), 256, 0, c->stream>>>(part, nblk, M, N, nb, alpha, beta, C, ldc);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// Workgroups of 256 threads of `kernel` resident on the device at once (CUs
// x occupancy): a grid of that size runs in one round.
static int64_t resident_blocks(const void* kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1024;
    }
    return (int64_t)(cus > 0 ? cus : 256) * (per > 0 ? per : 1);
}

// Runs the Gram kernel with RB x NB accumulators per thread: one contiguous
// run of k per workgroup, as many workgroups as are resident at once (no
// partly filled second round); the partial tiles go to the workspace.
template <int RB, int NB>
static int nn_gram_go(xr_ctx* c, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, int64_t* nblk_out, double** part_out) {
    static const int64_t res = resident_blocks(reinterpret_cast<const void*>(&k_dgemm_nn_gram<RB, NB>));
    int64_t nblk = ceil_div(K, 2048);
    if (nblk > res) nblk = res;
    const int64_t rows = ceil_div(K, nblk);
    nblk = ceil_div(K, rows);
    double* part = (double*)xr_workspace(c, 52, nblk * 64 * NB * sizeof(double));
    if (!part) return xr_fail(c, XR_ENOMEM, "workspace (Gram)");
    k_dgemm_nn_gram<RB, NB><<<(unsigned)nblk, 256, 0, c->stream>>>(M, N, K, A, lda, B, ldb, rows, part);
    *nblk_out = nblk;
    *part_out = part;
    return XR_OK;
}

static int launch_nn_gram(xr_ctx* c, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
    const int nb = N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : 8));
    const int64_t q = ceil_div(M, 4);             // rows per wave
    const int rb = q <= 1 ? 1 : (q <= 2 ? 2 : (q <= 4 ? 4 : (q <= 8 ? 8 : 16)));
    ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
    prof.shape(M, N, K, 1, 0, 0);
    int64_t nblk = 0;
    double* part = nullptr;
    int rc;
#define XR_GRAM_GO(RB_, NB_) rc = nn_gram_go<RB_, NB_>(c, M, N, K, A, lda, B, ldb, &nblk, &part)
    if (nb == 1) {
        if (rb == 1) XR_GRAM_GO(1, 1);
        else if (rb == 2) XR_GRAM_GO(2, 1);
        else if (rb == 4) XR_GRAM_GO(4, 1);
        else if (rb == 8) XR_GRAM_GO(8, 1);
        else XR_GRAM_GO(16, 1);
    } else if (nb == 2) {
        if (rb == 1) XR_GRAM_GO(1, 2);
        else if (rb == 2) XR_GRAM_GO(2, 2);
        else if (rb == 4) XR_GRAM_GO(4, 2);
        else XR_GRAM_GO(8, 2);
    } else if (nb == 4) {
        if (rb == 1) XR_GRAM_GO(1, 4);
        else if (rb == 2) XR_GRAM_GO(2, 4);
        else XR_GRAM_GO(4, 4);
    } else {
        if (rb == 1) XR_GRAM_GO(1, 8);
        else XR_GRAM_GO(2, 8);
    }
#undef XR_GRAM_GO
    if (rc) return rc;
    XR_LAUNCH_CHECK(c);
    k_tn_reduce<<<(unsigned)(M * N), 256, 0, c->stream>>>(part, nblk, M, N, nb, alpha, beta, C, ldc);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// K split for `units` output tiles on `slots` resident workgroups:
// returns the best score (ideal time / modelled time) and the split.  Rounds
// of resident workgroups each take K/s + 32 (per-workgroup start-up and
// store, in k units); the reduction reads and writes (s + 1) M N doubles
// against the product's 2 M N K flops (~46 flops per byte at the kernel's
// rate).  Each split keeps at least kmin of K: 256 in general, 64 for a
// narrow output (one row or column of tiles), whose partials are few bytes
// next to the operand it streams.
static double split_choice(int64_t units, int64_t slots, int64_t K, int64_t* splits, int64_t kmin = 256) {
    auto score = [&](int64_t s) {
        const double rounds = std::ceil((double)(units * s) / (double)slots);
        const double ideal = (double)units * (double)K / (double)slots;
        const double t = rounds * ((double)K / (double)s + 32.0);
        return ideal / t / (1.0 + (s > 1 ? (double)(s + 1) * 46.0 / (double)K : 0.0));
    };
    double best = score(1);
    *splits = 1;
    if (K < 512 || units >= 4 * slots) return best;
    // up to two rounds of resident workgroups (at most 512 splits): a few
    // output tiles over a long K are bound by how many CUs stream operands
    const int64_t smax = std::min<int64_t>(std::min<int64_t>(512, std::max<int64_t>(64, 2 * slots / units)), K / kmin);
    for (int64_t s = 2; s <= smax; s++) {
        const double e = score(s);
        if (e > best) { best = e; *splits = s; }
    }
    return best;
}

template <int BM, int BN>
static int64_t tile_slots() {
    static const int64_t r = resident_blocks(reinterpret_cast<const void*>(&k_dgemm<BM, BN, false, false>));
    return r;
}

template <int BM, int BN>
static int launch_tiles(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                        const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                        int64_t sB, double beta, double* C, int64_t ldc, int64_t sC,
                        int64_t batch, int splits_req = 0, bool upper = false) {
    int64_t tm = ceil_div(M, BM), tn = ceil_div(N, BN);
    int64_t tiles = tm * tn;
    if (tiles > 0x7fffffff || batch > 65535)
        return xr_fail(c, XR_EUNSUPPORTED, "GEMM grid too large");
    // split K when the output tiles leave the last round of resident
    // workgroups partly empty; the partial products go to scratch and are
    // summed in a fixed order (deterministic)
    int64_t splits = 1, kchunk = K;
    if (splits_req > 0) {                    // from the autotune table
        kchunk = ceil_div(ceil_div(K, splits_req), GEMM_BK) * GEMM_BK;
        if (kchunk < GEMM_BK) kchunk = GEMM_BK;
        splits = ceil_div(K, kchunk);
        if (splits < 1) { splits = 1; kchunk = K; }
    } else {
        split_choice(tiles * batch, tile_slots<BM, BN>(), K, &splits, (tm == 1 || tn == 1) ? 64 : 256);
        if (splits > 1) {
            kchunk = ceil_div(ceil_div(K, splits), GEMM_BK) * GEMM_BK;
            splits = ceil_div(K, kchunk);
        }
    }
    // accumulate in the epilogue when the tiles fit in one round of half the
    // resident workgroups (the variant's lower occupancy then costs nothing),
    // or when K is short: the product is then bound by the C traffic, and the
    // staged path's workspace write and reduction pass would double it
    if (!upper && splits == 1 && beta != 0.0 && (2 * tiles * batch <= tile_slots<BM, BN>() || K <= 256)) {
        dim3 grid((unsigned)tiles, 1, (unsigned)batch);
        ProfScope prof(c, XR_K_GEMM_F64, 2.0 * M * N * K * batch, 8.0 * (M * K + K * N + 2 * M * N) * batch);
        prof.shape(M, N, K, batch, ta, tb);
#define XR_GEMM_ACC(TA_, TB_)                                                                    \
    k_dgemm_acc<BM, BN, TA_, TB_><<<grid, GEMM_THREADS, 0, c->stream>>>(M, N, K, alpha, A, lda, sA, \
                                                                        B, ldb, sB, beta, C, ldc,  \
                                                                        sC, tm, tn)
        if (ta && tb) XR_GEMM_ACC(true, true);
        else if (ta) XR_GEMM_ACC(true, false);
        else if (tb) XR_GEMM_ACC(false, true);
        else XR_GEMM_ACC(false, false);
#undef XR_GEMM_ACC
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    double* Cw = C;
    int64_t ldw = ldc, sCw = sC, sSplit = 0;
    double aw = alpha, bw = beta;
    const bool staged = splits > 1 || beta != 0.0;
    if (staged) {
        Cw = (double*)xr_workspace(c, 50, splits * batch * M * N * sizeof(double));
        if (!Cw) return xr_fail(c, XR_ENOMEM, "workspace (split-K)");
        ldw = N;
        sCw = M * N;
        sSplit = batch * M * N;
        aw = 1.0;
        bw = 0.0;
    }
    dim3 grid((unsigned)tiles, (unsigned)splits, (unsigned)batch);
    {
        ProfScope prof(c, XR_K_GEMM_F64, 2.0 * M * N * K * batch, 8.0 * (M * K + K * N + M * N) * batch);
        prof.shape(M, N, K, batch, ta, tb);
#define XR_GEMM_LAUNCH(TA_, TB_)                                                                 \
    k_dgemm<BM, BN, TA_, TB_><<<grid, GEMM_THREADS, 0, c->stream>>>(M, N, K, aw, A, lda, sA, B,   \
                                                                     ldb, sB, bw, Cw, ldw, sCw,   \
                                                                     tm, tn, kchunk, sSplit, upper ? 1 : 0)
        if (ta && tb) XR_GEMM_LAUNCH(true, true);
        else if (ta) XR_GEMM_LAUNCH(true, false);
        else if (tb) XR_GEMM_LAUNCH(false, true);
        else XR_GEMM_LAUNCH(false, false);
#undef XR_GEMM_LAUNCH
        XR_LAUNCH_CHECK(c);
    }
    if (staged) {
        ProfScope prof(c, XR_K_OTHER, 0.0, 8.0 * (splits + 2) * batch * M * N);
        int rc = splitk_reduce(c, Cw, splits, batch, M, N, alpha, beta, C, ldc, sC);
        if (rc) return rc;
    }
    return XR_OK;
}

// C[i][j] = C[j][i] for i > j: the lower triangle from the upper one (32 x 32
// tiles through LDS; a workgroup writes one lower tile and reads its mirror,
// which no workgroup writes).
__global__ __launch_bounds__(256) void k_mirror_upper(double* __restrict__ C, int64_t n, int64_t ldc) {
    const int64_t a = blockIdx.y, b = blockIdx.x;     // target tile rows a, columns b
    if (a < b) return;
    __shared__ double T[32][33];
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int j = ty; j < 32; j += 8) {                // T[j][tx] = C[b*32 + j][a*32 + tx]
        const int64_t r = b * 32 + j, col = a * 32 + tx;
        if (r < n && col < n) T[j][tx] = C[r * ldc + col];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {                // C[a*32 + j][b*32 + tx] = C[b*32 + tx][a*32 + j]
        const int64_t r = a * 32 + j, col = b * 32 + tx;
        if (r < n && col < n && r > col) C[r * ldc + col] = T[tx][j];
    }
}

// C = alpha A^T A for A (K x n, leading dimension lda): the tile GEMM on the
// upper tiles only, then the lower triangle mirrored.
int xr_dgemm_syrk_tn(xr_ctx* c, int64_t n, int64_t K, double alpha, const double* A, int64_t lda, double* C,
                     int64_t ldc) {
    if (n <= 0) return XR_OK;
    int rc = launch_tiles<128, 128>(c, 1, 0, n, n, K, alpha, A, lda, 0, A, lda, 0, 0.0, C, ldc, 0, 1, 0, true);
    if (rc) return rc;
    const int64_t t = ceil_div(n, 32);
    if (t > 65535) return xr_fail(c, XR_EUNSUPPORTED, "matrix too large for the mirror grid");
    ProfScope prof(c, XR_K_OTHER, 0.0, 8.0 * n * n);
    k_mirror_upper<<<dim3((unsigned)t, (unsigned)t), dim3(32, 8), 0, c->stream>>>(C, n, ldc);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

// C = alpha A A^T for A (n x K, leading dimension lda): the same upper-tile
// product and mirror.
int xr_dgemm_syrk_nt(xr_ctx* c, int64_t n, int64_t K, double alpha, const double* A, int64_t lda, double* C,
                     int64_t ldc) {
    if (n <= 0) return XR_OK;
    int rc = launch_tiles<128, 128>(c, 0, 1, n, n, K, alpha, A, lda, 0, A, lda, 0, 0.0, C, ldc, 0, 1, 0, true);
    if (rc) return rc;
    const int64_t t = ceil_div(n, 32);
    if (t > 65535) return xr_fail(c, XR_EUNSUPPORTED, "matrix too large for the mirror grid");
    ProfScope prof(c, XR_K_OTHER, 0.0, 8.0 * n * n);
    k_mirror_upper<<<dim3((unsigned)t, (unsigned)t), dim3(32, 8), 0, c->stream>>>(C, n, ldc);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}

extern "C" int xr_gemm_syrk(xr_ctx* c, int trans, int64_t n, int64_t K, double alpha, const double* A, int64_t lda,
                            double* C, int64_t ldc) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n < 0 || K < 0) return xr_fail(c, XR_EINVAL, "negative size");
    if (ldc < n) return xr_fail(c, XR_ESHAPE, "ldc < n");
    if ((trans ? lda < n : lda < K) && n > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "lda too small");
    if (n == 0) return XR_OK;
    return trans ? xr_dgemm_syrk_tn(c, n, K, alpha, A, lda, C, ldc) : xr_dgemm_syrk_nt(c, n, K, alpha, A, lda, C, ldc);
}

// One explicit configuration: tile 0 = 128x128, 1 = 128x64, 2 = 64x128,
// 3 = 64x64; splits = K splits (0 = the heuristic).
static int launch_cfg(xr_ctx* c, int tile, int splits, int ta, int tb, int64_t M, int64_t N, int64_t K,
                      double alpha, const double* A, int64_t lda, int64_t sA, const double* B,
                      int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc, int64_t sC,
                      int64_t batch) {
    switch (tile) {
    case 0: return launch_tiles<128, 128>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, splits);
    case 1: return launch_tiles<128, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, splits);
    case 2: return launch_tiles<64, 128>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, splits);
    default: return launch_tiles<64, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, splits);
    }
}

// Chooses the tile shape: the autotuned configuration when the shape has
// one, else 128x128 when the grid still fills the chip and smaller tiles
// for skinny or small problems.
int xr_dgemm_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                    const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                    int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return XR_OK;
    if (ta && !tb && batch == 1 && M <= 64 && N <= 8 && K >= 4096)
        return launch_tn_skinny(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (ta && batch == 1 && M >= 1024 && K <= 64 && N <= 8) {
        ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
        prof.shape(M, N, K, batch, ta, tb);
#define XR_TN_TALL(NB_)                                                                                 \
        if (tb) k_dgemm_tn_tall<NB_, true><<<xr_grid(M), 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc); \
        else k_dgemm_tn_tall<NB_, false><<<xr_grid(M), 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)
        if (N <= 1) { XR_TN_TALL(1); }
        else if (N <= 2) { XR_TN_TALL(2); }
        else if (N <= 4) { XR_TN_TALL(4); }
        else { XR_TN_TALL(8); }
#undef XR_TN_TALL
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    if (!ta && !tb && batch == 1 && N <= 8 && K >= 4096 && M * (N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : 8))) <= 64)
        return launch_nn_gram(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (N <= 8 && !ta && batch == 1 && M >= 1024 && K <= 4) {
        ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
        prof.shape(M, N, K, batch, ta, tb);
#define XR_ROWS_SMALL(NB_)                                                                              \
        if (tb) k_dgemm_rows_small<NB_, true><<<xr_grid(M), 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc); \
        else k_dgemm_rows_small<NB_, false><<<xr_grid(M), 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)
        if (N <= 1) { XR_ROWS_SMALL(1); }
        else if (N <= 2) { XR_ROWS_SMALL(2); }
        else if (N <= 4) { XR_ROWS_SMALL(4); }
        else { XR_ROWS_SMALL(8); }
#undef XR_ROWS_SMALL
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    if (N <= 8 && !ta && batch == 1 && M >= 1024 && K <= 64) {
        ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
        prof.shape(M, N, K, batch, ta, tb);
        const unsigned rgrid = (unsigned)(ceil_div(M, 64) < (1 << 20) ? ceil_div(M, 64) : (1 << 20));
#define XR_ROWWISE(NB_)                                                                                 \
        if (tb) k_dgemm_rows_lds<NB_, true><<<rgrid, 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc); \
        else k_dgemm_rows_lds<NB_, false><<<rgrid, 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)
        if (N <= 1) { XR_ROWWISE(1); }
        else if (N <= 2) { XR_ROWWISE(2); }
        else if (N <= 4) { XR_ROWWISE(4); }
        else { XR_ROWWISE(8); }
#undef XR_ROWWISE
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    if (N <= 8 && !ta && batch == 1 && M >= 1024 && K > 64) {
        ProfScope prof(c, XR_K_GEMM_BW, 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N));
        prof.shape(M, N, K, batch, ta, tb);
        unsigned grid = xr_grid(ceil_div(M, 4) * 64);
#define XR_SKINNY(NB_)                                                                                  \
        if (tb) k_dgemm_skinny<NB_, true><<<grid, 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc); \
        else k_dgemm_skinny<NB_, false><<<grid, 256, 0, c->stream>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)
        if (N <= 1) { XR_SKINNY(1); }
        else if (N <= 2) { XR_SKINNY(2); }
        else if (N <= 4) { XR_SKINNY(4); }
        else { XR_SKINNY(8); }
#undef XR_SKINNY
        XR_LAUNCH_CHECK(c);
        return XR_OK;
    }
    auto it = c->tune.find(std::array<int64_t, 6>{ta, tb, M, N, K, batch});
    if (it != c->tune.end())
        return launch_cfg(c, it->second.first, it->second.second, ta, tb, M, N, K, alpha, A, lda, sA,
                          B, ldb, sB, beta, C, ldc, sC, batch);
    // a short-K update of C is a read-modify-write of C: the accumulating
    // 64 x 64 tiles keep more workgroups (and C loads) in flight
    if (beta != 0.0 && K <= 64)
        return launch_tiles<64, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    // 128 x 128 tiles when they (with a K split where K allows one) fill the
    // rounds of resident workgroups well; smaller tiles for skinny or
    // small problems
    int64_t big = ceil_div(M, 128) * ceil_div(N, 128) * batch, s128 = 1;
    if (big >= 512 || (M > 64 && N > 64 && split_choice(big, tile_slots<128, 128>(), K, &s128) >= 0.8))
        return launch_tiles<128, 128>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    if (big < 64)                  // small products: the most (and K-split) workgroups
        return launch_tiles<64, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    if (N <= 64 && M > 64)
        return launch_tiles<128, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    if (M <= 64 && N > 64)
        return launch_tiles<64, 128>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    return launch_tiles<64, 64>(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

extern "C" int xr_gemm_f64(xr_ctx* c, int transA, int transB, int64_t M, int64_t N, int64_t K,
                           double alpha, const double* A, int64_t lda, int64_t sA,
                           const double* B, int64_t ldb, int64_t sB, double beta, double* C,
                           int64_t ldc, int64_t sC, int64_t batch) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M < 0 || N < 0 || K < 0 || batch < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
    if (ldc < N) return xr_fail(c, XR_ESHAPE, "ldc < N");
    if ((transA ? lda < M : lda < K) && M > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "lda too small");
    if ((transB ? ldb < K : ldb < N) && N > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "ldb too small");
    return xr_dgemm_launch(c, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                           sC, batch);
}

// Launches the problems of one tile shape, XR_GROUP_PROBS per grid (with
// beta != 0 the accumulating kernel, C_q = alpha op(A_q) op(B_q) + beta C_q;
// lower: square products with BM == BN whose grids hold only the tiles on
// and below the diagonal; kind: the profiler class of the launches).
template <int BM, int BN>
static int launch_grouped(xr_ctx* c, const std::vector<int>& ids, int ta, int tb, const int64_t* M,
                          const int64_t* N, const int64_t* K, double alpha, const double* const* A,
                          const int64_t* lda, const double* const* B, const int64_t* ldb, double* const* C,
                          const int64_t* ldc, double beta = 0.0, int lower = 0, int kind = XR_K_GEMM_GROUPED) {
    if (lower && BM != BN) return xr_fail(c, XR_EUNSUPPORTED, "lower-triangle grids need square tiles");
    // launches of equal problem counts, problems by decreasing K
    std::vector<int> order(ids);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return K[a] > K[b]; });
    const size_t nl = (order.size() + XR_GROUP_PROBS - 1) / XR_GROUP_PROBS;
    for (size_t l = 0; l < nl; l++) {
        GemmGroup g;
        g.n = 0;
        g.lower = lower;
        g.alpha = alpha;
        g.beta = beta;
        int64_t slots = 0;                         // workgroups per XCD
        double flops = 0.0, bytes = 0.0;
        for (size_t t = l; t < order.size(); t += nl) {
            const int q = order[t];
            GemmProb& P = g.p[g.n++];
            P = GemmProb{A[q], B[q], C[q], M[q], N[q], K[q], lda[q], ldb[q], ldc[q],
                         ceil_div(M[q], BM), ceil_div(N[q], BN)};
            if (lower && M[q] != N[q]) return xr_fail(c, XR_EUNSUPPORTED, "lower-triangle grid of a non-square C");
            slots += ceil_div(lower ? P.tiles_m * (P.tiles_m + 1) / 2 : P.tiles_m * P.tiles_n, 8);
            flops += (lower ? 1.0 : 2.0) * M[q] * N[q] * K[q];
            bytes += 8.0 * (M[q] * K[q] + K[q] * N[q] + M[q] * N[q]);
        }
        if (8 * slots > 0x7fffffff) return xr_fail(c, XR_EUNSUPPORTED, "GEMM grid too large");
        ProfScope prof(c, kind, flops, bytes);
        prof.shape(g.p[0].M, g.p[0].N, g.p[0].K, g.n, ta, tb);
        const dim3 grid((unsigned)(8 * slots));
        if (beta != 0.0) {
            if (ta || !tb) return xr_fail(c, XR_EUNSUPPORTED, "accumulating grouped GEMM: A B^T only");
            k_dgemm_grouped<BM, BN, false, true, true><<<grid, GEMM_THREADS, 0, c->stream>>>(g);
        } else if (ta && tb) k_dgemm_grouped<BM, BN, true, true><<<grid, GEMM_THREADS, 0, c->stream>>>(g);
        else if (ta) k_dgemm_grouped<BM, BN, true, false><<<grid, GEMM_THREADS, 0, c->stream>>>(g);
        else if (tb) k_dgemm_grouped<BM, BN, false, true><<<grid, GEMM_THREADS, 0, c->stream>>>(g);
        else k_dgemm_grouped<BM, BN, false, false><<<grid, GEMM_THREADS, 0, c->stream>>>(g);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

// C_q = alpha op(A_q) op(B_q) for n independent problems (row-major, the
// same transposes).  Problems with N > 64 share 128 x 128 tile grids, those
// with 8 < N <= 64 share 128 x 64 grids (both with M > 64); the rest (the
// tall-skinny and small shapes) go to the single-product dispatch, whose
// bandwidth kernels suit them.  No K split and no beta: every C_q is
// overwritten.
extern "C" int xr_gemm_grouped(xr_ctx* c, int n, int ta, int tb, const int64_t* M, const int64_t* N,
                               const int64_t* K, double alpha, const double* const* A, const int64_t* lda,
                               const double* const* B, const int64_t* ldb, double* const* C,
                               const int64_t* ldc) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(c->mu);
    if (n < 0) return xr_fail(c, XR_EINVAL, "negative problem count");
    std::vector<int> wide, narrow;
    for (int q = 0; q < n; q++) {
        if (M[q] < 0 || N[q] < 0 || K[q] < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
        if (ldc[q] < N[q]) return xr_fail(c, XR_ESHAPE, "ldc < N");
        if ((ta ? lda[q] < M[q] : lda[q] < K[q]) && M[q] > 0 && K[q] > 0)
            return xr_fail(c, XR_ESHAPE, "lda too small");
        if ((tb ? ldb[q] < K[q] : ldb[q] < N[q]) && N[q] > 0 && K[q] > 0)
            return xr_fail(c, XR_ESHAPE, "ldb too small");
    }
    for (int q = 0; q < n; q++) {
        if (M[q] == 0 || N[q] == 0) continue;
        if (M[q] > 64 && N[q] > 64) wide.push_back(q);
        else if (M[q] > 64 && N[q] > 8) narrow.push_back(q);
        else {
            const int rc = xr_dgemm_launch(c, ta, tb, M[q], N[q], K[q], alpha, A[q], lda[q], 0, B[q], ldb[q], 0,
                                           0.0, C[q], ldc[q], 0, 1);
            if (rc) return rc;
        }
    }
    int rc = launch_grouped<128, 128>(c, wide, ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc);
    if (rc) return rc;
    return launch_grouped<128, 64>(c, narrow, ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc);
}

// ---------------------------------------------------------------------------
// Autotuner: times every tile shape with a ladder of K splits on scratch
// operands of the given shape (HIP events on the context stream) and keeps
// the fastest in the context's table, which xr_dgemm_launch consults.
// ---------------------------------------------------------------------------
extern "C" int xr_gemm_tune(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, int64_t batch,
                            int reps, int* tile_out, int* splits_out, double* ms_out) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || reps <= 0) return xr_fail(c, XR_EINVAL, "bad tune shape");
    if ((N <= 8 && !ta && batch == 1 && M >= 1024) ||
        (ta && !tb && batch == 1 && M <= 64 && N <= 8 && K >= 4096) ||
        (ta && batch == 1 && M >= 1024 && K <= 64 && N <= 8) ||
        (!ta && !tb && batch == 1 && N <= 8 && K >= 4096 &&
         M * (N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : 8))) <= 64)) {   // bandwidth kernels, not tiled
        if (tile_out) *tile_out = -1;
        if (splits_out) *splits_out = 0;
        if (ms_out) *ms_out = 0.0;
        return XR_OK;
    }
    const int64_t na = M * K * batch, nb = K * N * batch, nc = M * N * batch;
    if ((na + nb + nc) * 8 > (int64_t)16 << 30) return xr_fail(c, XR_EUNSUPPORTED, "shape too large to tune");
    double* A = (double*)xr_workspace(c, 60, na * sizeof(double));
    double* B = (double*)xr_workspace(c, 61, nb * sizeof(double));
    double* C = (double*)xr_workspace(c, 62, nc * sizeof(double));
    if (!A || !B || !C) return xr_fail(c, XR_ENOMEM, "workspace (tune)");
    int rc;
    if ((rc = xr_fill(c, A, na, 1e-3)) || (rc = xr_fill(c, B, nb, 1e-3))) return rc;
    const int64_t lda = ta ? M : K, ldb = tb ? K : N;
    const int64_t sA = batch > 1 ? M * K : 0, sB = batch > 1 ? K * N : 0, sC = batch > 1 ? M * N : 0;
    hipEvent_t e0, e1;
    XR_HIP(c, hipEventCreate(&e0));
    XR_HIP(c, hipEventCreate(&e1));
    const bool prof = c->prof;
    c->prof = false;
    const std::array<int64_t, 6> key{ta, tb, M, N, K, batch};
    c->tune.erase(key);
    // the untuned dispatch is the first candidate: a table entry is kept
    // only for a configuration that beats it
    double best = 1e300;
    int best_tile = -1, best_split = 0;
    rc = xr_dgemm_launch(c, ta, tb, M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, C, N, sC, batch);
    if (!rc) {
        XR_HIP(c, hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; r++)
            rc |= xr_dgemm_launch(c, ta, tb, M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, C, N, sC, batch);
        XR_HIP(c, hipEventRecord(e1, c->stream));
        XR_HIP(c, hipEventSynchronize(e1));
        float ms = 0.f;
        XR_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        best = ms / reps;
    }
    for (int tile = 0; tile < 4 && !rc; tile++) {
        for (int sp = 1; sp <= 64; sp *= 2) {
            if (sp > 1 && K / sp < 256) break;
            rc = launch_cfg(c, tile, sp, ta, tb, M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, C, N, sC, batch);
            if (rc) break;
            XR_HIP(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; r++)
                rc |= launch_cfg(c, tile, sp, ta, tb, M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, C, N, sC, batch);
            XR_HIP(c, hipEventRecord(e1, c->stream));
            XR_HIP(c, hipEventSynchronize(e1));
            float ms = 0.f;
            XR_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            if (rc) break;
            if (ms / reps < best) { best = ms / reps; best_tile = tile; best_split = sp; }
        }
        if (rc) break;
    }
    c->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_tile >= 0) c->tune[key] = std::make_pair(best_tile, best_split);
    if (tile_out) *tile_out = best_tile;
    if (splits_out) *splits_out = best_split;
    if (ms_out) *ms_out = best;
    return XR_OK;
}

extern "C" int xr_gemm_tune_reset(xr_ctx* c) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->tune.clear();
    return XR_OK;
}

// ---------------------------------------------------------------------------
// Split-FP32 GEMM on v_mfma_f32_16x16x4f32 with FP64 accumulation.
//
// While an operand tile is staged in LDS, every group of 8 consecutive k
// values of one row of A (one column of B) is put on a common power-of-two
// grid g = 2^(e-10), e = exponent of the group's largest |x|:
//   hi = rint(x / g) * g      (|hi / g| <= 2^10: 11 significant bits)
//   lo = fp32(x - hi)         (|lo| <= g/2, 24 significant bits)
// A hi*hi product is an integer multiple of ga*gb below 2^20 of it, so the
// 8 products of one group (two chained MFMAs) sum to < 2^23 grid units and
// the FP32 matrix core forms them EXACTLY; that sum is then added into the
// FP64 accumulator.  The remaining terms hi_a*lo_b + lo_a*(hi_b + lo_b) are
// ~2^-11 of the group's scale and go through FP32 accumulation over one
// 16-deep LDS tile; their rounding, and the 24-bit lo, leave a relative
// error of a few 2^-35 of max|a|*max|b| per group.
// ---------------------------------------------------------------------------
typedef float f4_t __attribute__((ext_vector_type(4)));

// Biased exponent field of |x|.
__device__ __forceinline__ int xr_expo(double x) {
    return (int)((__double_as_longlong(x) >> 52) & 0x7ff);
}

// Split x on the grid of a group whose largest biased exponent is eb.  A
// group whose largest value is below 2^-1012 (or zero) is flushed to zero.
__device__ __forceinline__ void xr_split(double x, int eb, float& hi, float& lo) {
    if (eb < 12) { hi = 0.f; lo = 0.f; return; }
    const double s = __longlong_as_double((long long)(2055 - eb) << 52);   // 2^(1032-eb) = 1/g
    const double gi = __longlong_as_double((long long)(eb - 9) << 52);     // 2^(eb-1032) = g
    const double h = rint(x * s) * gi;
    hi = (float)h;
    lo = (float)(x - h);
}

// Max over the 8 lanes (aligned) that hold one group.
__device__ __forceinline__ int xr_group8_max(int e) {
    e = max(e, __shfl_xor(e, 1));
    e = max(e, __shfl_xor(e, 2));
    e = max(e, __shfl_xor(e, 4));
    return e;
}

// BM x BM tiles on 4 * BM threads (4 elements of each operand per thread
// and 16-deep k step): 128 x 128 as 8 waves of 64 x 32, 64 x 64 as 4 waves
// of 32 x 32.
template <bool TA, bool TB, int BM>
__global__ __launch_bounds__(4 * BM) void k_dgemm_split(
    int64_t M, int64_t N, int64_t K, const double* __restrict__ A, int64_t lda, int64_t sA,
    const double* __restrict__ B, int64_t ldb, int64_t sB, double* __restrict__ C, int64_t ldc,
    int64_t sC, int64_t tiles_m, int64_t tiles_n, int64_t kchunk, int64_t sSplit) {
    constexpr int THREADS = 4 * BM, BN = BM, PAD = 4;
    constexpr int WGN = (BM == 128) ? 4 : 2, WGM = THREADS / 64 / WGN;
    constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
    constexpr int PER = BM * GEMM_BK / THREADS;
    static_assert(PER == 4 && (BM == 128 || BM == 64), "4 elements per thread and operand");
    __shared__ float Ah[GEMM_BK][BM + PAD], Al[GEMM_BK][BM + PAD];
    __shared__ float Bh[GEMM_BK][BN + PAD], Bl[GEMM_BK][BN + PAD];
    const int64_t batch = blockIdx.z;
    A += batch * sA;
    B += batch * sB;
    C += batch * sC + (int64_t)blockIdx.y * sSplit;
    const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
    K = (K - kbeg < kchunk) ? (K - kbeg) : kchunk;
    if (TA) A += kbeg * lda; else A += kbeg;
    if (TB) B += kbeg; else B += kbeg * ldb;
    const int64_t tile = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    const int64_t GROUP = 8, per_group = GROUP * tiles_n, g = tile / per_group, first_m = g * GROUP;
    const int64_t gsize = (tiles_m - first_m < GROUP) ? (tiles_m - first_m) : GROUP;
    const int64_t tm = first_m + (tile % per_group) % gsize, tn = (tile % per_group) / gsize;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave / WGN, wn = wave % WGN;
    // Element i of a thread as (row, k) inside a BM x 16 tile.  Operands
    // stored k-major (A transposed, B plain) give lanes t and t^1 the two
    // halves of one row's 8-deep group; row-major ones give 8 aligned lanes
    // one k each of a row's group.
    auto pos = [&](bool kmajor, int i, int& row, int& kk) {
        if (kmajor) { row = (tid >> 1) & (BM - 1); kk = (tid / (2 * BM)) * 8 + (tid & 1) * 4 + i; }
        else { kk = tid % GEMM_BK; row = tid / GEMM_BK + i * (THREADS / GEMM_BK); }
    };
    auto group_max = [&](bool kmajor, int* e) {
        if (kmajor) {
            int m = max(max(e[0], e[1]), max(e[2], e[3]));
            m = max(m, __shfl_xor(m, 1));
#pragma unroll
            for (int i = 0; i < PER; i++) e[i] = m;
        } else {
#pragma unroll
            for (int i = 0; i < PER; i++) e[i] = xr_group8_max(e[i]);
        }
    };
    double ra[PER], rb[PER];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            int mm, kk;
            pos(TA, i, mm, kk);
            int64_t gm = m0 + mm, gk = k0 + kk;
            ra[i] = (gm < M && gk < K) ? (TA ? A[gk * lda + gm] : A[gm * lda + gk]) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            int nn, kk;
            pos(!TB, i, nn, kk);
            int64_t gn = n0 + nn, gk = k0 + kk;
            rb[i] = (gn < N && gk < K) ? (TB ? B[gn * ldb + gk] : B[gk * ldb + gn]) : 0.0;
        }
    };
    auto store = [&]() {
        int ea[PER], eb[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) { ea[i] = xr_expo(ra[i]); eb[i] = xr_expo(rb[i]); }
        group_max(TA, ea);
        group_max(!TB, eb);
#pragma unroll
        for (int i = 0; i < PER; i++) {
            int mm, kk;
            pos(TA, i, mm, kk);
            xr_split(ra[i], ea[i], Ah[kk][mm], Al[kk][mm]);
            pos(!TB, i, mm, kk);
            xr_split(rb[i], eb[i], Bh[kk][mm], Bl[kk][mm]);
        }
    };
    d4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    const int64_t nk = (K + GEMM_BK - 1) / GEMM_BK;
    const int fr = lane & 15, fk = lane >> 4;
    if (nk > 0) { load(0); store(); __syncthreads(); }
    const f4_t z4 = {0.f, 0.f, 0.f, 0.f};
    for (int64_t kt = 0; kt < nk; kt++) {
        if (kt + 1 < nk) load((kt + 1) * GEMM_BK);
#pragma unroll
        for (int grp = 0; grp < GEMM_BK / 8; grp++) {
            float ah[2][FM], al[2][FM], bh[2][FN], bl[2][FN], bs[2][FN];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int k = grp * 8 + s * 4 + fk;
#pragma unroll
                for (int i = 0; i < FM; i++) {
                    ah[s][i] = Ah[k][wm * WM + i * 16 + fr];
                    al[s][i] = Al[k][wm * WM + i * 16 + fr];
                }
#pragma unroll
                for (int j = 0; j < FN; j++) {
                    bh[s][j] = Bh[k][wn * WN + j * 16 + fr];
                    bl[s][j] = Bl[k][wn * WN + j * 16 + fr];
                    bs[s][j] = bh[s][j] + bl[s][j];
                }
            }
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++) {
                    f4_t hh = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[0][i], bh[0][j], z4, 0, 0, 0);
                    hh = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[1][i], bh[1][j], hh, 0, 0, 0);
                    f4_t x = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[0][i], bl[0][j], z4, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x4f32(al[0][i], bs[0][j], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[1][i], bl[1][j], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x4f32(al[1][i], bs[1][j], x, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[i][j][r] += (double)hh[r] + (double)x[r];
                }
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        if (kt + 1 < nk) { store(); __syncthreads(); }
    }
    // D[row][col] of v_mfma_f32_16x16x4f32: col = lane % 16, row = XR_F32_ROW(fk, r)
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int64_t gm = m0 + wm * WM + i * 16 + XR_F32_ROW(fk, r);
                int64_t gn = n0 + wn * WN + j * 16 + fr;
                if (gm < M && gn < N) C[gm * ldc + gn] = acc[i][j][r];
            }
}

template <int BM>
static void split_go(xr_ctx* c, int ta, int tb, dim3 grid, int64_t M, int64_t N, int64_t K, const double* A,
                     int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB, double* Cw,
                     int64_t tm, int64_t tn, int64_t kchunk, int64_t batch) {
    const int nt = 4 * BM;
    if (ta && tb) k_dgemm_split<true, true, BM><<<grid, nt, 0, c->stream>>>(M, N, K, A, lda, sA, B, ldb, sB, Cw, N, M * N, tm, tn, kchunk, batch * M * N);
    else if (ta) k_dgemm_split<true, false, BM><<<grid, nt, 0, c->stream>>>(M, N, K, A, lda, sA, B, ldb, sB, Cw, N, M * N, tm, tn, kchunk, batch * M * N);
    else if (tb) k_dgemm_split<false, true, BM><<<grid, nt, 0, c->stream>>>(M, N, K, A, lda, sA, B, ldb, sB, Cw, N, M * N, tm, tn, kchunk, batch * M * N);
    else k_dgemm_split<false, false, BM><<<grid, nt, 0, c->stream>>>(M, N, K, A, lda, sA, B, ldb, sB, Cw, N, M * N, tm, tn, kchunk, batch * M * N);
}

// Split-FP32 product (always staged: partial tiles go to a workspace and the
// fixed-order reduction applies beta).
int xr_dgemm_split_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                          const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                          int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return XR_OK;
    // 64 x 64 tiles, with a K split (chunks of at least 256) toward 1024
    // workgroups when the tiles are few.  The 128 x 128 variant needs more
    // than the 256 registers a lane has at two waves per SIMD (it spilled 27)
    // and ran the c3 split iteration 6% slower.
    const int bm = 64;
    int64_t tm = ceil_div(M, bm), tn = ceil_div(N, bm), tiles = tm * tn;
    if (tiles > 0x7fffffff || batch > 65535) return xr_fail(c, XR_EUNSUPPORTED, "GEMM grid too large");
    int64_t splits = 1, kchunk = K;
    if (tiles * batch < 512 && K >= 512) {
        splits = ceil_div(1024, tiles * batch);
        int64_t max_split = K / 256;
        if (splits > max_split) splits = max_split;
        if (splits > 512) splits = 512;
        if (splits < 1) splits = 1;
        kchunk = ceil_div(ceil_div(K, splits), GEMM_BK) * GEMM_BK;
        splits = ceil_div(K, kchunk);
    }
    double* Cw = (double*)xr_workspace(c, 51, splits * batch * M * N * sizeof(double));
    if (!Cw) return xr_fail(c, XR_ENOMEM, "workspace (split-FP32)");
    dim3 grid((unsigned)tiles, (unsigned)splits, (unsigned)batch);
    {
        ProfScope prof(c, XR_K_GEMM_SPLIT, 2.0 * M * N * K * batch, 8.0 * (M * K + K * N + M * N) * batch);
        prof.shape(M, N, K, batch, ta, tb);
        split_go<64>(c, ta, tb, grid, M, N, K, A, lda, sA, B, ldb, sB, Cw, tm, tn, kchunk, batch);
        XR_LAUNCH_CHECK(c);
    }
    return splitk_reduce(c, Cw, splits, batch, M, N, alpha, beta, C, ldc, sC);
}

// The contraction GEMMs of the two-electron terms go through here: FP64,
// split-FP32 or split-int8 depending on the context's precision setting
// (split-int8 for the products where it pays, xr_i8_pays; FP64 for the rest).
int xr_gemm_2e(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
               const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB,
               double beta, double* C, int64_t ldc, int64_t sC, int64_t batch) {
    if (c->precision == XR_PREC_SPLIT_F32)
        return xr_dgemm_split_launch(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                                     sC, batch);
    if (c->precision == XR_PREC_SPLIT_I8 && xr_i8_pays(M, N, K, batch))
        return xr_dgemm_i8_launch(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    return xr_dgemm_launch(c, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

extern "C" int xr_set_precision(xr_ctx* c, int mode) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (mode != XR_PREC_F64 && mode != XR_PREC_SPLIT_F32 && mode != XR_PREC_SPLIT_I8)
        return xr_fail(c, XR_EINVAL, "unknown precision mode");
    c->precision = mode;
    return XR_OK;
}

extern "C" int xr_gemm_split_f32(xr_ctx* c, int transA, int transB, int64_t M, int64_t N, int64_t K,
                                 double alpha, const double* A, int64_t lda, int64_t sA,
                                 const double* B, int64_t ldb, int64_t sB, double beta, double* C,
                                 int64_t ldc, int64_t sC, int64_t batch) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M < 0 || N < 0 || K < 0 || batch < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
    if (ldc < N) return xr_fail(c, XR_ESHAPE, "ldc < N");
    return xr_dgemm_split_launch(c, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                                 sC, batch);
}

// ---------------------------------------------------------------------------
// No-allocation FP64 GEMM
// ---------------------------------------------------------------------------
// Staging memory of xr_gemm_f64_noalloc (the split-K partial tiles and the
// partial sums of the tall-skinny kernels): one 64 MiB array per device,
// held until the process ends and shared by every such call on that device.
// Host code never passes the array's name to a kernel (in host code it is a
// host-side stand-in); it uses the device address xr_noalloc_init stores in
// the context.
#define XR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) double scratch[XR_NOALLOC_SCRATCH_BYTES / sizeof(double)];

// Called by xr_create with the context's device current: the device address
// of this device's copy of `scratch`.
int xr_noalloc_init(xr_ctx* c) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return xr_fail(c, XR_EHIP, "hipGetSymbolAddress(scratch) failed");
    }
    c->scratch = p;
    c->scratch_bytes = XR_NOALLOC_SCRATCH_BYTES;
    return XR_OK;
}

extern "C" int xr_gemm_f64_noalloc(xr_ctx* c, int transA, int transB, int64_t M, int64_t N, int64_t K,
                                   double alpha, const double* A, int64_t lda, int64_t sA,
                                   const double* B, int64_t ldb, int64_t sB, double beta, double* C,
                                   int64_t ldc, int64_t sC, int64_t batch) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (M < 0 || N < 0 || K < 0 || batch < 0) return xr_fail(c, XR_EINVAL, "negative GEMM size");
    if (ldc < N) return xr_fail(c, XR_ESHAPE, "ldc < N");
    if ((transA ? lda < M : lda < K) && M > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "lda too small");
    if ((transB ? ldb < K : ldb < N) && N > 0 && K > 0) return xr_fail(c, XR_ESHAPE, "ldb too small");
    if (!c->scratch) return xr_fail(c, XR_EHIP, "no scratch address for this device");
    // the same kernels and choices as xr_gemm_f64, with every staging buffer
    // taken from the scratch array; the launch timers are off for the call
    // (they would create events)
    const bool prof = c->prof;
    c->prof = false;
    c->noalloc = true;
    c->noalloc_blocked = false;
    int rc = xr_dgemm_launch(c, transA, transB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                             batch);
    c->noalloc = false;
    c->noalloc_blocked = false;
    c->prof = prof;
    return rc;
}
