// Batched pair blocks of an XR build, and graph capture without a framework.
//
// build_matrix_elements.dimer (general-XRCC/build_H.py:56-102) for every
// state pair of every charge combination ("job") of every fragment pair,
// through the dimer families of general-XRCC/H_contractions.c:95-180, as
// one program of a fixed handful of launches instead of 2-4 GEMMs, permutes
// and scatters per job:
//
//   moves    zero the outputs, permute each distinct integral block once,
//            copy the density columns the one-electron-transfer product
//            concatenates                                         1 launch
//   stage 1  density x permuted-integral products of every job: work-list
//            FP64 MFMA GEMM, one launch per (transpose, tile) class, long
//            K split over workgroups                          <= 6 launches
//   reduce   split and multi-term partial products, summed in fixed order
//   stage 2  the bilinear forms' final products (A B^T)       <= 3 launches
//   reduce
//   scatter  every job's block, with its spectator-diagonal vectors and
//            nuclear constant, into the outputs                   1 launch
//   symm     A + A^T of every output                              1 launch
//
// The descriptors (problems, tile tables, moves, reductions, scatters) are
// device tables.  A plan owns them and its intermediates, so running it
// allocates nothing and issues the same launches on the same addresses,
// which xr_graph_begin / xr_graph_end capture into a HIP graph.
#include "xr_common.h"
#include "xr_gemm.h"
#include <vector>
#include <map>
#include <tuple>
#include <algorithm>
#include <cstring>

// ---------------------------------------------------------------------------
// device descriptors
// ---------------------------------------------------------------------------
struct WlGemm {
    const double* A;
    const double* B;
    double* C;
    int64_t M, N, K, lda, ldb, ldc;
    double alpha;
    int64_t kchunk;        // K range of one split
    int64_t split_stride;  // C offset between the outputs of consecutive splits
};
struct WlTile {
    int32_t prob, split, tm, tn;   // prob < 0: padding (uneven XCD shares)
};
struct WlMove {                     // dst[o . dst_st] = src[o . src_st] over dims (zero when src is null)
    const double* src;
    double* dst;
    int64_t dims[4];
    int64_t src_st[4];
    int64_t dst_st[4];
    int64_t count;
};
struct WlReduce {                   // dst = sum of nparts consecutive M x N partials
    const double* src;
    double* dst;
    int64_t M, N, ldd, nparts;
};
struct WlScatter {
    const double* H;                // element (a, b) at H[a * sa + b * sb]
    int64_t sa, sb;
    const double* e1;               // [a], added where I2 == J2 (or null)
    const double* e2;               // [b], added where I1 == J1 (or null)
    double scale, nuc;
    double* out;
    int64_t ldo, mul;
    const int64_t* r1;
    const int64_t* c1;
    const int64_t* r2;
    const int64_t* c2;
    int64_t nJ1, nJ2, nA, nB;
    int64_t tiles_b, tile0;         // 32 x 32 tiles over (a, b); first tile id of the job
};
struct WlSym {
    double* A;
    int64_t n, ld, t, tile0;        // t = ceil(n / 32); t * t tiles from tile0
};

#define PAIR_MOVE_BLOCKS 1024
#define PAIR_GRID_MAX (1 << 20)

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pair_moves(const WlMove* __restrict__ mv) {
    const WlMove m = mv[blockIdx.y];
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m.dims[0] == 1 && m.dims[1] == 1 && m.dims[2] == 1) {      // one axis
        const int64_t ss = m.src_st[3], ds = m.dst_st[3];
        if (m.src) for (; i < m.count; i += step) m.dst[i * ds] = m.src[i * ss];
        else for (; i < m.count; i += step) m.dst[i * ds] = 0.0;
        return;
    }
    for (; i < m.count; i += step) {
        int64_t x = i, so = 0, d = 0;
#pragma unroll
        for (int j = 3; j >= 0; j--) {
            const int64_t o = x % m.dims[j];
            x /= m.dims[j];
            so += o * m.src_st[j];
            d += o * m.dst_st[j];
        }
        m.dst[d] = m.src ? m.src[so] : 0.0;
    }
}

// One output tile per workgroup from the work list: problem, K split and
// tile coordinates come from the table (XCD-aware order built on the host).
template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(GEMM_THREADS) void k_dgemm_worklist(const WlGemm* __restrict__ probs,
                                                                 const WlTile* __restrict__ tiles) {
    const WlTile t = tiles[blockIdx.x];
    if (t.prob < 0) return;
    const WlGemm P = probs[t.prob];
    const int64_t kbeg = (int64_t)t.split * P.kchunk;
    const int64_t K = (P.K - kbeg < P.kchunk) ? (P.K - kbeg) : P.kchunk;
    const double* A = P.A + (TA ? kbeg * P.lda : kbeg);
    const double* B = P.B + (TB ? kbeg : kbeg * P.ldb);
    double* C = P.C + (int64_t)t.split * P.split_stride;
    dgemm_tile<BM, BN, TA, TB>(P.M, P.N, K, P.alpha, A, P.lda, B, P.ldb, C, P.ldc, (int64_t)t.tm * BM,
                               (int64_t)t.tn * BN);
}

__global__ __launch_bounds__(256) void k_pair_reduce(const WlReduce* __restrict__ rd) {
    const WlReduce r = rd[blockIdx.y];
    const int64_t n = r.M * r.N;
    XR_GRID_STRIDE(i, n) {
        double acc = 0.0;
        for (int64_t s = 0; s < r.nparts; s++) acc += r.src[s * n + i];
        const int64_t row = i / r.N, col = i - row * r.N;
        r.dst[row * r.ldd + col] = acc;
    }
}

// 32 x 32 (a, b) tiles of every job: the block tile goes through LDS (read
// along whichever index is contiguous in H), the placement of the tile's 32
// a and 32 b values is computed once, and the writes run along b (J2, the
// contiguous output column).
__global__ __launch_bounds__(256) void k_pair_scatter(const WlScatter* __restrict__ js, int n, int64_t total) {
    __shared__ double T[32][33];
    __shared__ int64_t rowA[32], colA[32], rowB[32], colB[32];
    __shared__ double eA[32], eB[32];
    __shared__ int diagA[32], diagB[32];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (js[mid].tile0 <= tile) lo = mid;
            else hi = mid - 1;
        }
        const WlScatter& J = js[lo];
        const int64_t local = tile - J.tile0;
        const int64_t a0 = (local / J.tiles_b) * 32, b0 = (local % J.tiles_b) * 32;
        if (ty == 0) {
            const int64_t a = a0 + tx;
            if (a < J.nA) {
                const int64_t I1 = a / J.nJ1, J1 = a - I1 * J.nJ1;
                rowA[tx] = J.r1[I1] * J.mul;
                colA[tx] = J.c1[J1] * J.mul;
                diagA[tx] = I1 == J1;
                eA[tx] = J.e1 ? J.e1[a] : 0.0;
            }
        } else if (ty == 1) {
            const int64_t b = b0 + tx;
            if (b < J.nB) {
                const int64_t I2 = b / J.nJ2, J2 = b - I2 * J.nJ2;
                rowB[tx] = J.r2[I2];
                colB[tx] = J.c2[J2];
                diagB[tx] = I2 == J2;
                eB[tx] = J.e2 ? J.e2[b] : 0.0;
            }
        }
        if (J.sb == 1) {
            for (int j = ty; j < 32; j += 8) {
                const int64_t a = a0 + j, b = b0 + tx;
                T[j][tx] = (a < J.nA && b < J.nB) ? J.H[a * J.sa + b] : 0.0;
            }
        } else {            // sa == 1: H is [b, a]
            for (int j = ty; j < 32; j += 8) {
                const int64_t b = b0 + j, a = a0 + tx;
                T[tx][j] = (a < J.nA && b < J.nB) ? J.H[b * J.sb + a] : 0.0;
            }
        }
        __syncthreads();
        for (int j = ty; j < 32; j += 8) {
            const int64_t a = a0 + j, b = b0 + tx;
            if (a < J.nA && b < J.nB) {
                double v = T[j][tx];
                if (diagA[j]) v += eB[tx];
                if (diagB[tx]) v += eA[j];
                if (diagA[j] && diagB[tx]) v += J.nuc;
                J.out[(rowA[j] + rowB[tx]) * J.ldo + colA[j] + colB[tx]] = J.scale * v;
            }
        }
        __syncthreads();
    }
}

// A = A + A^T for every output: workgroup tile (p, q) with p <= q holds the
// 32 x 32 tiles (p, q) and (q, p) in LDS and writes both sums.
__global__ __launch_bounds__(256) void k_pair_symm(const WlSym* __restrict__ ss, int n, int64_t total) {
    __shared__ double P[32][33], Q[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ss[mid].tile0 <= tile) lo = mid;
            else hi = mid - 1;
        }
        const WlSym& S = ss[lo];
        const int64_t local = tile - S.tile0;
        const int64_t p = local / S.t, q = local - p * S.t;
        if (p > q) continue;                       // uniform over the workgroup
        const int64_t rp = p * 32, rq = q * 32, N = S.n, ld = S.ld;
        double* A = S.A;
        for (int j = ty; j < 32; j += 8) {
            const int64_t r1 = rp + j, c1 = rq + tx, r2 = rq + j, c2 = rp + tx;
            P[j][tx] = (r1 < N && c1 < N) ? A[r1 * ld + c1] : 0.0;
            Q[j][tx] = (r2 < N && c2 < N) ? A[r2 * ld + c2] : 0.0;
        }
        __syncthreads();
        for (int j = ty; j < 32; j += 8) {
            const int64_t r1 = rp + j, c1 = rq + tx, r2 = rq + j, c2 = rp + tx;
            if (r1 < N && c1 < N) A[r1 * ld + c1] = P[j][tx] + Q[tx][j];
            if (p != q && r2 < N && c2 < N) A[r2 * ld + c2] = Q[j][tx] + P[tx][j];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// program construction (host)
// ---------------------------------------------------------------------------
namespace {

struct Proto {                      // one product before the K-split decision
    int ta, tb;
    int64_t M, N, K;
    double alpha;
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    int group;                      // output group (several terms may add into one output)
    bool two_e;                     // a two-electron product (a split GEMM in the split modes)
};
struct Group {
    double* dst;
    int64_t M, N, ldd;
    std::vector<int> terms;
};

struct GemmLaunch {
    int ta, tb, shape;              // shape 0: 128 x 128, 1: 128 x 64, 2: 64 x 64
    int64_t tile_off, ntiles;
    double flops, bytes;            // of the launch's products (profiler records)
};
struct SplitCall {                  // a product run by its own split launch (FP32 or int8)
    WlGemm g;
    int ta, tb, mode;
};

struct Program {
    std::vector<WlGemm> probs;
    std::vector<WlTile> tiles;
    std::vector<GemmLaunch> launches[2];
    std::vector<SplitCall> split_calls[2];
    std::vector<WlMove> moves;
    std::vector<WlReduce> reduces[2];
    std::vector<WlScatter> scat;
    int64_t scat_tiles = 0;
    std::vector<WlSym> syms;
    int64_t sym_tiles = 0;
    size_t arena_bytes = 0;
    double flops = 0.0;
    int64_t workgroups = 0;
    // device table offsets (bytes) in the descriptor blob
    size_t o_probs = 0, o_tiles = 0, o_moves = 0, o_red[2] = {0, 0}, o_scat = 0, o_sym = 0, blob_bytes = 0;
};

const int64_t SHAPE_BM[3] = {128, 128, 64}, SHAPE_BN[3] = {128, 64, 64};

template <int BM, int BN, bool TA, bool TB>
int64_t worklist_slots() {
    static const int64_t r = resident_blocks(reinterpret_cast<const void*>(&k_dgemm_worklist<BM, BN, TA, TB>));
    return r;
}

int64_t class_slots(int ta, int tb, int shape) {
    (void)ta;
    // resident workgroups depend on the tile shape (LDS, registers), not on
    // the transposes
    switch (shape) {
    case 0: return tb ? worklist_slots<128, 128, false, true>() : worklist_slots<128, 128, false, false>();
    case 1: return tb ? worklist_slots<128, 64, false, true>() : worklist_slots<128, 64, false, false>();
    default: return tb ? worklist_slots<64, 64, false, true>() : worklist_slots<64, 64, false, false>();
    }
}

int shape_of(int64_t M, int64_t N) {
    if (M > 64 && N > 64) return 0;
    if (M > 64) return 1;
    return 2;
}

void tile_coords(int64_t local, int64_t tm_n, int64_t tn_n, int64_t* tm, int64_t* tn) {
    const int64_t GROUP = 8, per = GROUP * tn_n;
    const int64_t g = local / per, first = g * GROUP;
    const int64_t gs = std::min<int64_t>(GROUP, tm_n - first);
    *tm = first + (local % per) % gs;
    *tn = (local % per) / gs;
}

struct Builder {
    uintptr_t base;                 // arena base (0 in the sizing pass)
    int mode;
    Program& P;
    std::string err;
    std::vector<Proto> protos[2];
    std::vector<Group> groups[2];
    std::map<std::tuple<const double*, int, int64_t, int64_t>, double*> wcache;

    Builder(uintptr_t b, int m, Program& p) : base(b), mode(m), P(p) {}

    double* take(int64_t count) {
        size_t off = (P.arena_bytes + 255) & ~(size_t)255;
        P.arena_bytes = off + (size_t)std::max<int64_t>(count, 1) * sizeof(double);
        return reinterpret_cast<double*>(base + off);
    }

    void move(const double* src, double* dst, std::array<int64_t, 4> dims, std::array<int64_t, 4> sst,
              std::array<int64_t, 4> dst_st) {
        WlMove m;
        m.src = src;
        m.dst = dst;
        m.count = 1;
        for (int j = 0; j < 4; j++) {
            m.dims[j] = dims[j];
            m.src_st[j] = sst[j];
            m.dst_st[j] = dst_st[j];
            m.count *= dims[j];
        }
        if (m.count > 0) P.moves.push_back(m);
    }

    // W = V (d0 x d1 x d2 x d3, row-major) with its axes permuted (output
    // axis j = input axis p[j]), flattened; one copy per distinct block.
    double* permuted(const double* V, int kind, int64_t d[4], const int p[4]) {
        auto key = std::make_tuple(V, kind, d[0], d[3]);
        auto it = wcache.find(key);
        if (it != wcache.end()) return it->second;
        const int64_t ist[4] = {d[1] * d[2] * d[3], d[2] * d[3], d[3], 1};
        std::array<int64_t, 4> od, os, ds;
        for (int j = 0; j < 4; j++) {
            od[j] = d[p[j]];
            os[j] = ist[p[j]];
        }
        ds = {od[1] * od[2] * od[3], od[2] * od[3], od[3], 1};
        double* W = take(d[0] * d[1] * d[2] * d[3]);
        move(V, W, od, os, ds);
        wcache[key] = W;
        return W;
    }

    int group(int stage, double* dst, int64_t M, int64_t N, int64_t ldd) {
        groups[stage].push_back(Group{dst, M, N, ldd, {}});
        return (int)groups[stage].size() - 1;
    }

    void product(int stage, int g, int ta, int tb, int64_t K, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, bool two_e) {
        Group& G = groups[stage][g];
        protos[stage].push_back(Proto{ta, tb, G.M, G.N, K, alpha, A, lda, B, ldb, g, two_e});
        G.terms.push_back((int)protos[stage].size() - 1);
    }

    int job(const xr_pair_job& j) {
        const int64_t nA = j.nI1 * j.nJ1, nB = j.nI2 * j.nJ2;
        const int64_t nAf = j.transposed ? nB : nA, nBf = j.transposed ? nA : nB;
        const int64_t n1 = j.n1, n2 = j.n2;
        if (nA == 0 || nB == 0) return XR_OK;
        double* H = nullptr;
        const double *e1 = nullptr, *e2 = nullptr;
        double nuc = 0.0;
        if (j.kind == XR_PAIR_2MIN2PLS) {
            int64_t d[4] = {n1, n1, n2, n2};
            const int p[4] = {0, 1, 3, 2};           // W[(p1 q1), (s2 r2)] = V[p1, q1, r2, s2]
            double* W = permuted(j.V[0], 0, d, p);
            double* G = take(nAf * n2 * n2);
            product(0, group(0, G, nAf, n2 * n2, n2 * n2), 0, 0, n1 * n1, 1.0, j.R1[0], n1 * n1, W, n2 * n2, true);
            H = take(nAf * nBf);
            product(1, group(1, H, nAf, nBf, nBf), 0, 1, n2 * n2, 1.0, G, n2 * n2, j.R2[0], n2 * n2, true);
        } else if (j.kind == XR_PAIR_EXEX) {
            int64_t d[4] = {n1, n2, n1, n2};
            const int p[4] = {0, 2, 1, 3};           // W[(p1 r1), (q2 s2)] = V[p1, q2, r1, s2]
            double* W = permuted(j.V[0], 1, d, p);
            double* G = take(nA * n2 * n2);
            product(0, group(0, G, nA, n2 * n2, n2 * n2), 0, 0, n1 * n1, 1.0, j.R1[0], n1 * n1, W, n2 * n2, true);
            double* v2 = take(nB);                   // e2[b] = sum Rca2[b, pq] U122[pq]
            product(0, group(0, v2, nB, 1, 1), 0, 0, n2 * n2, 1.0, j.R2[0], n2 * n2, j.h[0], 1, false);
            double* v1 = take(nA);                   // e1[a] = sum Rca1[a, pq] U211[pq]
            product(0, group(0, v1, nA, 1, 1), 0, 0, n1 * n1, 1.0, j.R1[0], n1 * n1, j.h[1], 1, false);
            e1 = v1;
            e2 = v2;
            nuc = j.nuc;
            H = take(nA * nB);
            product(1, group(1, H, nA, nB, nB), 0, 1, n2 * n2, 4.0, G, n2 * n2, j.R2[0], n2 * n2, true);
        } else {                                     // XR_PAIR_1MIN1PLS
            const int64_t n1c = n1 * n1 * n1, n2c = n2 * n2 * n2, Kc = n1 + n2;
            int64_t d1[4] = {n1, n1, n1, n2};
            const int p1[4] = {1, 0, 2, 3};          // W1[(q1 p1 r1), s2] = V1112[p1, q1, r1, s2]
            double* W1 = permuted(j.V[0], 2, d1, p1);
            int64_t d2[4] = {n1, n2, n2, n2};
            const int p2[4] = {0, 1, 3, 2};          // W2[p1, (q2 s2 r2)] = V1222[p1, q2, r2, s2]
            double* W2 = permuted(j.V[1], 3, d2, p2);
            double* Ac = take(nAf * Kc);
            double* Bc = take(nBf * Kc);
            // Ac = [2 Rcca1 W1 + Rc1 h | Rc1],  Bc = [Ra2 | 2 Rcaa2 W2^T],  H = Ac Bc^T
            move(j.R1[0], Ac + n2, {1, 1, nAf, n1}, {0, 0, n1, 1}, {0, 0, Kc, 1});
            move(j.R2[0], Bc, {1, 1, nBf, n2}, {0, 0, n2, 1}, {0, 0, Kc, 1});
            const int ga = group(0, Ac, nAf, n2, Kc);
            product(0, ga, 0, 0, n1c, 2.0, j.R1[1], n1c, W1, n2, true);
            product(0, ga, 0, 0, n1, 1.0, j.R1[0], n1, j.h[0], n2, false);
            product(0, group(0, Bc + n2, nBf, n1, Kc), 0, 1, n2c, 2.0, j.R2[1], n2c, W2, n2c, true);
            H = take(nAf * nBf);
            product(1, group(1, H, nAf, nBf, nBf), 0, 1, Kc, 1.0, Ac, Kc, Bc, Kc, true);
        }
        WlScatter s;
        s.H = H;
        s.sa = j.transposed ? 1 : nB;
        s.sb = j.transposed ? nA : 1;
        s.e1 = e1;
        s.e2 = e2;
        s.scale = j.scale;
        s.nuc = nuc;
        s.out = j.out;
        s.ldo = j.ldo;
        s.mul = j.mul;
        s.r1 = j.rows1;
        s.c1 = j.cols1;
        s.r2 = j.rows2;
        s.c2 = j.cols2;
        s.nJ1 = j.nJ1;
        s.nJ2 = j.nJ2;
        s.nA = nA;
        s.nB = nB;
        s.tiles_b = ceil_div(nB, 32);
        s.tile0 = P.scat_tiles;
        P.scat_tiles += ceil_div(nA, 32) * s.tiles_b;
        P.scat.push_back(s);
        return XR_OK;
    }

    // K splits, partial buffers, reductions and the XCD-aware tile tables of
    // one stage's products.
    void finish_stage(int st) {
        std::vector<Proto>& pr = protos[st];
        std::vector<int64_t> splits(pr.size(), 1);
        // two-electron products that run as their own split launches: all of
        // them in split-FP32, those where int8 pays in split-int8
        std::vector<char> sep(pr.size(), 0);
        for (size_t q = 0; q < pr.size(); q++)
            sep[q] = pr[q].two_e && (mode == XR_PREC_SPLIT_F32 ||
                                     (mode == XR_PREC_SPLIT_I8 && xr_i8_pays(pr[q].M, pr[q].N, pr[q].K, 1)));
        // per (ta, tb, shape) class: balance the work over the resident workgroups
        std::map<std::array<int, 3>, std::vector<int>> classes;
        for (size_t q = 0; q < pr.size(); q++) {
            if (sep[q]) continue;
            classes[{pr[q].ta, pr[q].tb, shape_of(pr[q].M, pr[q].N)}].push_back((int)q);
        }
        for (auto& kv : classes) {
            const int shape = kv.first[2];
            const int64_t bm = SHAPE_BM[shape], bn = SHAPE_BN[shape];
            double work = 0.0;
            for (int q : kv.second) work += (double)ceil_div(pr[q].M, bm) * ceil_div(pr[q].N, bn) * pr[q].K;
            const double slots = (double)class_slots(kv.first[0], kv.first[1], shape);
            // tiles of at most half the balanced per-workgroup K work (and K
            // chunks of at least 256) keep any one tile from running long
            const double target = std::max(256.0, 0.5 * work / slots);
            for (int q : kv.second) {
                int64_t s = (int64_t)std::ceil((double)pr[q].K / target);
                s = std::max<int64_t>(1, std::min<int64_t>(s, std::min<int64_t>(256, pr[q].K / 256)));
                splits[q] = s;
            }
        }
        // K chunks in whole LDS steps (the rounded chunk may need fewer splits)
        std::vector<int64_t> kchunk(pr.size(), 0);
        for (size_t q = 0; q < pr.size(); q++) {
            int64_t s = splits[q], kc = pr[q].K;
            if (s > 1) {
                kc = ceil_div(ceil_div(pr[q].K, s), GEMM_BK) * GEMM_BK;
                s = ceil_div(pr[q].K, kc);
            }
            splits[q] = s;
            kchunk[q] = kc;
            P.flops += 2.0 * pr[q].M * pr[q].N * pr[q].K;
        }
        // outputs: direct for one unsplit term, else partial slots + a reduction
        std::vector<double*> dst(pr.size(), nullptr);
        std::vector<int64_t> ldc(pr.size(), 0), sstride(pr.size(), 0);
        for (Group& G : groups[st]) {
            int64_t slots = 0;
            for (int q : G.terms) slots += splits[q];
            if (G.terms.size() == 1 && slots == 1) {
                const int q = G.terms[0];
                dst[q] = G.dst;
                ldc[q] = G.ldd;
                continue;
            }
            double* part = take(slots * G.M * G.N);
            int64_t s0 = 0;
            for (int q : G.terms) {
                dst[q] = part + s0 * G.M * G.N;
                ldc[q] = G.N;
                sstride[q] = G.M * G.N;
                s0 += splits[q];
            }
            P.reduces[st].push_back(WlReduce{part, G.dst, G.M, G.N, G.ldd, slots});
        }
        // tile tables: by class, problems by decreasing K chunk, each
        // problem's tiles in contiguous runs over the 8 XCDs
        for (auto& kv : classes) {
            const int ta = kv.first[0], tb = kv.first[1], shape = kv.first[2];
            const int64_t bm = SHAPE_BM[shape], bn = SHAPE_BN[shape];
            std::vector<int> order = kv.second;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return kchunk[a] > kchunk[b]; });
            std::vector<std::vector<WlTile>> xl(8);
            int rot = 0;
            double lfl = 0.0, lby = 0.0;
            for (int q : order) {
                const Proto& p = pr[q];
                lfl += 2.0 * p.M * p.N * p.K;
                lby += 8.0 * (p.M * p.K + p.K * p.N + p.M * p.N);
                WlGemm g{p.A, p.B, dst[q], p.M, p.N, p.K, p.lda, p.ldb, ldc[q], p.alpha, kchunk[q], sstride[q]};
                const int id = (int)P.probs.size();
                P.probs.push_back(g);
                const int64_t tm_n = ceil_div(p.M, bm), tn_n = ceil_div(p.N, bn), per = tm_n * tn_n;
                const int64_t total = per * splits[q];
                const int64_t b8 = total / 8, r8 = total % 8;
                int64_t next = 0;
                for (int k = 0; k < 8; k++) {
                    const int x = (rot + k) & 7;
                    const int64_t share = b8 + (k < r8 ? 1 : 0);
                    for (int64_t u = 0; u < share; u++, next++) {
                        const int64_t sp = next / per, local = next - sp * per;
                        int64_t tm, tn;
                        tile_coords(local, tm_n, tn_n, &tm, &tn);
                        xl[x].push_back(WlTile{id, (int32_t)sp, (int32_t)tm, (int32_t)tn});
                    }
                }
                rot = (rot + (int)r8) & 7;
            }
            size_t len = 0;
            for (auto& l : xl) len = std::max(len, l.size());
            GemmLaunch L{ta, tb, shape, (int64_t)P.tiles.size(), (int64_t)(8 * len), lfl, lby};
            for (size_t j = 0; j < len; j++)
                for (int x = 0; x < 8; x++)
                    P.tiles.push_back(j < xl[x].size() ? xl[x][j] : WlTile{-1, 0, 0, 0});
            P.launches[st].push_back(L);
            P.workgroups += L.ntiles;
        }
        for (size_t q = 0; q < pr.size(); q++)
            if (sep[q]) {
                const Proto& p = pr[q];
                WlGemm g{p.A, p.B, dst[q], p.M, p.N, p.K, p.lda, p.ldb, ldc[q], p.alpha, p.K, 0};
                P.split_calls[st].push_back(SplitCall{g, p.ta, p.tb, mode});
            }
    }
};

int check_job(const xr_pair_job& j, std::string& err) {
    if (j.kind < XR_PAIR_2MIN2PLS || j.kind > XR_PAIR_EXEX) { err = "unknown pair job kind"; return XR_EINVAL; }
    if (j.n1 <= 0 || j.n2 <= 0 || j.n1 > XR_MAX_SO || j.n2 > XR_MAX_SO) { err = "pair job spin orbitals out of range"; return XR_EINVAL; }
    if (j.nI1 < 0 || j.nJ1 < 0 || j.nI2 < 0 || j.nJ2 < 0) { err = "negative state counts"; return XR_EINVAL; }
    if (j.kind == XR_PAIR_EXEX && (j.transposed || j.nI1 != j.nJ1 || j.nI2 != j.nJ2)) {
        err = "exciton-exciton jobs keep both charges (square state blocks, not transposed)";
        return XR_EINVAL;
    }
    if (j.nI1 * j.nJ1 == 0 || j.nI2 * j.nJ2 == 0) return XR_OK;
    const int need_r = j.kind == XR_PAIR_EXEX ? 1 : (j.kind == XR_PAIR_1MIN1PLS ? 2 : 1);
    for (int k = 0; k < need_r; k++)
        if (!j.R1[k] || !j.R2[k]) { err = "pair job without its densities"; return XR_EINVAL; }
    if (!j.V[0] || (j.kind == XR_PAIR_1MIN1PLS && !j.V[1])) { err = "pair job without its integral blocks"; return XR_EINVAL; }
    if ((j.kind == XR_PAIR_1MIN1PLS && !j.h[0]) || (j.kind == XR_PAIR_EXEX && (!j.h[0] || !j.h[1]))) {
        err = "pair job without its one-electron blocks";
        return XR_EINVAL;
    }
    if (!j.out || !j.rows1 || !j.cols1 || !j.rows2 || !j.cols2 || j.ldo <= 0 || j.mul < 0) {
        err = "pair job without its placement";
        return XR_EINVAL;
    }
    return XR_OK;
}

int build_program(const xr_pair_job* jobs, int n_jobs, double* const* outs, const int64_t* out_n,
                  const int64_t* out_ld, int n_outs, int mode, uintptr_t base, Program& P, std::string& err) {
    P = Program();
    Builder b(base, mode, P);
    for (int i = 0; i < n_outs; i++) {
        const int64_t n = out_n[i], ld = out_ld[i];
        if (n < 0 || ld < n || (n > 0 && !outs[i])) { err = "bad pair output"; return XR_EINVAL; }
        if (n == 0) continue;
        if (ld == n) b.move(nullptr, outs[i], {1, 1, 1, n * n}, {0, 0, 0, 1}, {0, 0, 0, 1});
        else b.move(nullptr, outs[i], {1, 1, n, n}, {0, 0, 0, 0}, {0, 0, ld, 1});
        const int64_t t = ceil_div(n, 32);
        P.syms.push_back(WlSym{outs[i], n, ld, t, P.sym_tiles});
        P.sym_tiles += t * t;
    }
    for (int i = 0; i < n_jobs; i++) {
        int rc = check_job(jobs[i], err);
        if (rc) return rc;
        if ((rc = b.job(jobs[i]))) return rc;
    }
    b.finish_stage(0);
    b.finish_stage(1);
    if (P.moves.size() > 65535 || P.reduces[0].size() > 65535 || P.reduces[1].size() > 65535) {
        err = "too many pair jobs for one program";
        return XR_EUNSUPPORTED;
    }
    for (int st = 0; st < 2; st++)
        for (const GemmLaunch& L : P.launches[st])
            if (L.ntiles > 0x7fffffffll) { err = "pair program tile table too large"; return XR_EUNSUPPORTED; }
    // descriptor blob layout
    size_t o = 0;
    auto place = [&](size_t bytes) { size_t at = (o + 255) & ~(size_t)255; o = at + bytes; return at; };
    P.o_probs = place(P.probs.size() * sizeof(WlGemm));
    P.o_tiles = place(P.tiles.size() * sizeof(WlTile));
    P.o_moves = place(P.moves.size() * sizeof(WlMove));
    P.o_red[0] = place(P.reduces[0].size() * sizeof(WlReduce));
    P.o_red[1] = place(P.reduces[1].size() * sizeof(WlReduce));
    P.o_scat = place(P.scat.size() * sizeof(WlScatter));
    P.o_sym = place(P.syms.size() * sizeof(WlSym));
    P.blob_bytes = o > 0 ? o : 256;
    return XR_OK;
}

void serialize(const Program& P, std::vector<char>& blob) {
    blob.assign(P.blob_bytes, 0);
    auto put = [&](size_t off, const void* src, size_t bytes) { if (bytes) std::memcpy(blob.data() + off, src, bytes); };
    put(P.o_probs, P.probs.data(), P.probs.size() * sizeof(WlGemm));
    put(P.o_tiles, P.tiles.data(), P.tiles.size() * sizeof(WlTile));
    put(P.o_moves, P.moves.data(), P.moves.size() * sizeof(WlMove));
    put(P.o_red[0], P.reduces[0].data(), P.reduces[0].size() * sizeof(WlReduce));
    put(P.o_red[1], P.reduces[1].data(), P.reduces[1].size() * sizeof(WlReduce));
    put(P.o_scat, P.scat.data(), P.scat.size() * sizeof(WlScatter));
    put(P.o_sym, P.syms.data(), P.syms.size() * sizeof(WlSym));
}

int64_t program_launches(const Program& P) {
    int64_t n = P.moves.empty() ? 0 : 1;
    for (int st = 0; st < 2; st++)
        n += (int64_t)P.launches[st].size() + (int64_t)P.split_calls[st].size() + (P.reduces[st].empty() ? 0 : 1);
    n += (P.scat.empty() ? 0 : 1) + (P.syms.empty() ? 0 : 1);
    return n;
}

int run_program(xr_ctx* c, const Program& P, const char* dblob) {
    const WlGemm* probs = reinterpret_cast<const WlGemm*>(dblob + P.o_probs);
    const WlTile* tiles = reinterpret_cast<const WlTile*>(dblob + P.o_tiles);
    if (!P.moves.empty()) {
        ProfScope prof(c, XR_K_OTHER, 0.0, 0.0);
        k_pair_moves<<<dim3(PAIR_MOVE_BLOCKS, (unsigned)P.moves.size()), 256, 0, c->stream>>>(
            reinterpret_cast<const WlMove*>(dblob + P.o_moves));
        XR_LAUNCH_CHECK(c);
    }
    for (int st = 0; st < 2; st++) {
        for (const GemmLaunch& L : P.launches[st]) {
            if (L.ntiles == 0) continue;
            ProfScope prof(c, XR_K_GEMM_WORKLIST, L.flops, L.bytes);
            const dim3 grid((unsigned)L.ntiles);
            const WlTile* t = tiles + L.tile_off;
#define XR_WL(BM_, BN_)                                                                                   \
            if (L.tb) k_dgemm_worklist<BM_, BN_, false, true><<<grid, GEMM_THREADS, 0, c->stream>>>(probs, t); \
            else k_dgemm_worklist<BM_, BN_, false, false><<<grid, GEMM_THREADS, 0, c->stream>>>(probs, t)
            if (L.shape == 0) { XR_WL(128, 128); }
            else if (L.shape == 1) { XR_WL(128, 64); }
            else { XR_WL(64, 64); }
#undef XR_WL
            XR_LAUNCH_CHECK(c);
        }
        for (const SplitCall& s : P.split_calls[st]) {
            const int rc = (s.mode == XR_PREC_SPLIT_I8)
                ? xr_dgemm_i8_launch(c, s.ta, s.tb, s.g.M, s.g.N, s.g.K, s.g.alpha, s.g.A, s.g.lda, 0, s.g.B,
                                     s.g.ldb, 0, 0.0, s.g.C, s.g.ldc, 0, 1)
                : xr_dgemm_split_launch(c, s.ta, s.tb, s.g.M, s.g.N, s.g.K, s.g.alpha, s.g.A, s.g.lda, 0,
                                        s.g.B, s.g.ldb, 0, 0.0, s.g.C, s.g.ldc, 0, 1);
            if (rc) return rc;
        }
        if (!P.reduces[st].empty()) {
            ProfScope prof(c, XR_K_OTHER, 0.0, 0.0);
            k_pair_reduce<<<dim3(PAIR_MOVE_BLOCKS, (unsigned)P.reduces[st].size()), 256, 0, c->stream>>>(
                reinterpret_cast<const WlReduce*>(dblob + P.o_red[st]));
            XR_LAUNCH_CHECK(c);
        }
    }
    if (!P.scat.empty() && P.scat_tiles > 0) {
        ProfScope prof(c, XR_K_SCATTER, 0.0, 0.0);
        const unsigned g = (unsigned)std::min<int64_t>(P.scat_tiles, PAIR_GRID_MAX);
        k_pair_scatter<<<g, 256, 0, c->stream>>>(reinterpret_cast<const WlScatter*>(dblob + P.o_scat),
                                                 (int)P.scat.size(), P.scat_tiles);
        XR_LAUNCH_CHECK(c);
    }
    if (!P.syms.empty() && P.sym_tiles > 0) {
        ProfScope prof(c, XR_K_SCATTER, 0.0, 0.0);
        const unsigned g = (unsigned)std::min<int64_t>(P.sym_tiles, PAIR_GRID_MAX);
        k_pair_symm<<<g, 256, 0, c->stream>>>(reinterpret_cast<const WlSym*>(dblob + P.o_sym),
                                              (int)P.syms.size(), P.sym_tiles);
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

enum { WS_PAIR_ARENA = 70, WS_PAIR_BLOB = 71 };

}  // namespace

struct xr_pair_plan {
    xr_ctx* c;
    Program P;
    void* arena = nullptr;
    void* blob = nullptr;
};

extern "C" int xr_pair_blocks(xr_ctx* c, const xr_pair_job* jobs, int n_jobs, double* const* outs,
                              const int64_t* out_n, const int64_t* out_ld, int n_outs, int mode) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(c->mu);
    if (n_jobs < 0 || n_outs < 0 || (n_jobs > 0 && !jobs) || (n_outs > 0 && (!outs || !out_n || !out_ld)))
        return xr_fail(c, XR_EINVAL, "bad pair job list");
    if (mode != XR_PREC_F64 && mode != XR_PREC_SPLIT_F32 && mode != XR_PREC_SPLIT_I8)
        return xr_fail(c, XR_EINVAL, "unknown precision mode");
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return xr_fail(c, XR_EINVAL, "xr_pair_blocks uploads its descriptors from host memory and cannot be "
                                     "captured; capture xr_pair_plan_run instead");
    Program P;
    std::string err;
    int rc = build_program(jobs, n_jobs, outs, out_n, out_ld, n_outs, mode, 0, P, err);
    if (rc) return xr_fail(c, rc, err);
    const size_t arena_bytes = P.arena_bytes;
    char* arena = (char*)xr_workspace(c, WS_PAIR_ARENA, arena_bytes);
    char* blob = (char*)xr_workspace(c, WS_PAIR_BLOB, P.blob_bytes);
    if (!arena || !blob) return xr_fail(c, XR_ENOMEM, "workspace (pair program)");
    if ((rc = build_program(jobs, n_jobs, outs, out_n, out_ld, n_outs, mode, (uintptr_t)arena, P, err)))
        return xr_fail(c, rc, err);
    std::vector<char> host;
    serialize(P, host);
    // descriptors go through a pinned staging buffer of the context; the
    // previous call's copy out of it must have finished before it is reused
    if (c->pair_copied) XR_HIP(c, hipEventSynchronize(c->pair_copied));
    else XR_HIP(c, hipEventCreateWithFlags(&c->pair_copied, hipEventDisableTiming));
    if (c->pair_host_bytes < host.size()) {
        if (c->pair_host) (void)hipHostFree(c->pair_host);
        c->pair_host = nullptr;
        c->pair_host_bytes = 0;
        XR_HIP(c, hipHostMalloc(&c->pair_host, host.size(), hipHostMallocDefault));
        c->pair_host_bytes = host.size();
    }
    std::memcpy(c->pair_host, host.data(), host.size());
    XR_HIP(c, hipMemcpyAsync(blob, c->pair_host, host.size(), hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, hipEventRecord(c->pair_copied, c->stream));
    return run_program(c, P, blob);
}

extern "C" int xr_pair_plan_create(xr_ctx* c, const xr_pair_job* jobs, int n_jobs, double* const* outs,
                                   const int64_t* out_n, const int64_t* out_ld, int n_outs, int mode,
                                   xr_pair_plan** plan) {
    if (!c || !plan) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(c->mu);
    *plan = nullptr;
    if (n_jobs < 0 || n_outs < 0 || (n_jobs > 0 && !jobs) || (n_outs > 0 && (!outs || !out_n || !out_ld)))
        return xr_fail(c, XR_EINVAL, "bad pair job list");
    if (mode != XR_PREC_F64 && mode != XR_PREC_SPLIT_F32 && mode != XR_PREC_SPLIT_I8)
        return xr_fail(c, XR_EINVAL, "unknown precision mode");
    xr_pair_plan* p = new xr_pair_plan();
    p->c = c;
    std::string err;
    int rc = build_program(jobs, n_jobs, outs, out_n, out_ld, n_outs, mode, 0, p->P, err);
    if (rc) {
        delete p;
        return xr_fail(c, rc, err);
    }
    const size_t ab = p->P.arena_bytes, bb = p->P.blob_bytes;
    if (hipMalloc(&p->arena, ab < 256 ? 256 : ab) != hipSuccess || hipMalloc(&p->blob, bb) != hipSuccess) {
        (void)hipGetLastError();
        if (p->arena) (void)hipFree(p->arena);
        if (p->blob) (void)hipFree(p->blob);
        delete p;
        return xr_fail(c, XR_ENOMEM, "pair plan memory");
    }
    if ((rc = build_program(jobs, n_jobs, outs, out_n, out_ld, n_outs, mode, (uintptr_t)p->arena, p->P, err))) {
        (void)hipFree(p->arena);
        (void)hipFree(p->blob);
        delete p;
        return xr_fail(c, rc, err);
    }
    std::vector<char> host;
    serialize(p->P, host);
    if (hipMemcpy(p->blob, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(p->arena);
        (void)hipFree(p->blob);
        delete p;
        return xr_fail(c, XR_EHIP, "pair plan upload");
    }
    *plan = p;
    return XR_OK;
}

extern "C" int xr_pair_plan_run(xr_pair_plan* p) {
    if (!p || !p->c) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(p->c->mu);
    return run_program(p->c, p->P, (const char*)p->blob);
}

extern "C" int xr_pair_plan_info(xr_pair_plan* p, int64_t* launches, int64_t* workgroups, double* flops,
                                 int64_t* arena_bytes) {
    if (!p) return XR_EINVAL;
    if (launches) *launches = program_launches(p->P);
    if (workgroups) *workgroups = p->P.workgroups;
    if (flops) *flops = p->P.flops;
    if (arena_bytes) *arena_bytes = (int64_t)p->P.arena_bytes;
    return XR_OK;
}

extern "C" int xr_pair_plan_destroy(xr_pair_plan* p) {
    if (!p) return XR_OK;
    if (p->c) (void)hipStreamSynchronize(p->c->stream);
    if (p->arena) (void)hipFree(p->arena);
    if (p->blob) (void)hipFree(p->blob);
    delete p;
    return XR_OK;
}

// ---------------------------------------------------------------------------
// graph capture of the context stream
// ---------------------------------------------------------------------------
struct xr_graph {
    xr_ctx* c;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int xr_graph_begin(xr_ctx* c) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(c->mu);
    if (!c->stream) return xr_fail(c, XR_EINVAL, "graph capture needs a context stream of its own (xr_set_stream)");
    if (c->prof) return xr_fail(c, XR_EINVAL, "switch the profiler off before a capture");
    c->pins++;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        c->pins--;
        return xr_hip_check(c, e, "hipStreamBeginCapture");
    }
    return XR_OK;
}

extern "C" int xr_graph_end(xr_ctx* c, xr_graph** out) {
    if (!c || !out) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(c->mu);
    *out = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || !g) {
        c->pins--;
        if (g) (void)hipGraphDestroy(g);
        return xr_hip_check(c, e != hipSuccess ? e : hipErrorUnknown, "hipStreamEndCapture");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        c->pins--;
        (void)hipGraphDestroy(g);
        return xr_hip_check(c, e, "hipGraphInstantiate");
    }
    xr_graph* r = new xr_graph();
    r->c = c;
    r->graph = g;
    r->exec = x;
    *out = r;
    return XR_OK;     // the pin stays with the graph
}

extern "C" int xr_graph_launch(xr_graph* g) {
    if (!g || !g->c) return XR_EINVAL;
    std::lock_guard<std::mutex> lock(g->c->mu);
    XR_HIP(g->c, hipGraphLaunch(g->exec, g->c->stream));
    return XR_OK;
}

extern "C" int xr_graph_destroy(xr_graph* g) {
    if (!g) return XR_OK;
    xr_ctx* c = g->c;
    if (c) {
        std::lock_guard<std::mutex> lock(c->mu);
        (void)hipStreamSynchronize(c->stream);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        c->pins--;
    }
    delete g;
    return XR_OK;
}
