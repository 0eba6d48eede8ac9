// Spin-orbital integral blocks from the spatial integrals, and their
// adjoint: the blocks the contractions read (the one-body blocks T and U,
// the two-body blocks V of fragment quadruples), cut from the (rotated)
// spatial h (n x n), U (n x n per nucleus) and V (n^4, chemist notation).
// Within a fragment the alpha spin orbitals come first.
//   one-body: out[(s, p), (s', q)] = delta(s, s') A[o1 + p, o2 + q]
//   two-body: out[(s0,p), (s1,q), (s2,r), (s3,s)] = 1/4 ( d(s0,s2) d(s1,s3) V[o0+p, o2+r, o1+q, o3+s]
//                                                       - d(s0,s3) d(s1,s2) V[o0+p, o3+s, o1+q, o2+r] )
// (V_pqrs = (1/4)(<pq|rs> - <pq|sr>) over spin orbitals).  One launch per
// block in both directions; the adjoint gathers, for every spatial element
// of a block's source region, the block gradient entries that read it, so
// each element has one writer per launch and the sums are in a fixed order.
// ---------------------------------------------------------------------------

__global__ void k_spin_1e(const double* __restrict__ A, int64_t n, int o1, int o2, int n1, int n2,
                          double* __restrict__ out) {
    const int64_t W = 2 * n2;
    XR_GRID_STRIDE(e, (int64_t)(2 * n1) * W) {
        const int a = (int)(e / W), b = (int)(e - (int64_t)a * W);
        const int sa = a >= n1, sb = b >= n2, p = a - sa * n1, q = b - sb * n2;
        out[e] = (sa == sb) ? A[(int64_t)(o1 + p) * n + o2 + q] : 0.0;
    }
}

__global__ void k_spin_1e_adj(const double* __restrict__ G, int64_t n, int o1, int o2, int n1, int n2,
                              double* __restrict__ dA) {
    XR_GRID_STRIDE(e, (int64_t)n1 * n2) {
        const int p = (int)(e / n2), q = (int)(e - (int64_t)p * n2);
        const int64_t W = 2 * n2;
        dA[(int64_t)(o1 + p) * n + o2 + q] += G[(int64_t)p * W + q] + G[(int64_t)(n1 + p) * W + n2 + q];
    }
}

__device__ __forceinline__ int64_t v4(int64_t n, int64_t i, int64_t j, int64_t k, int64_t l) {
    return ((i * n + j) * n + k) * n + l;
}

__global__ void k_spin_2e(const double* __restrict__ V, int64_t n, int4 o, int4 d, double* __restrict__ out) {
    const int64_t D1 = 2 * d.y, D2 = 2 * d.z, D3 = 2 * d.w;
    XR_GRID_STRIDE(e, (int64_t)(2 * d.x) * D1 * D2 * D3) {
        int64_t t = e;
        const int c3 = (int)(t % D3); t /= D3;
        const int c2 = (int)(t % D2); t /= D2;
        const int c1 = (int)(t % D1);
        const int c0 = (int)(t / D1);
        const int s0 = c0 >= d.x, s1 = c1 >= d.y, s2 = c2 >= d.z, s3 = c3 >= d.w;
        const int p = c0 - s0 * d.x, q = c1 - s1 * d.y, r = c2 - s2 * d.z, s = c3 - s3 * d.w;
        double v = 0.0;
        if (s0 == s2 && s1 == s3) v += V[v4(n, o.x + p, o.z + r, o.y + q, o.w + s)];
        if (s0 == s3 && s1 == s2) v -= V[v4(n, o.x + p, o.w + s, o.y + q, o.z + r)];
        out[e] = 0.25 * v;
    }
}

// Adjoint of the first (direct) term over its source region
// V[o0 + p, o2 + r, o1 + q, o3 + s]: dV += 1/4 sum_(a, b) G[(a,p), (b,q), (a,r), (b,s)].
__global__ void k_spin_2e_adj_direct(const double* __restrict__ G, int64_t n, int4 o, int4 d,
                                     double* __restrict__ dV) {
    const int64_t D1 = 2 * d.y, D2 = 2 * d.z, D3 = 2 * d.w;
    XR_GRID_STRIDE(e, (int64_t)d.x * d.z * d.y * d.w) {
        int64_t t = e;                      // region order (p, r, q, s)
        const int s = (int)(t % d.w); t /= d.w;
        const int q = (int)(t % d.y); t /= d.y;
        const int r = (int)(t % d.z);
        const int p = (int)(t / d.z);
        double g = 0.0;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                g += G[((int64_t)(a * d.x + p) * D1 + b * d.y + q) * D2 * D3 + (int64_t)(a * d.z + r) * D3 + b * d.w + s];
        dV[v4(n, o.x + p, o.z + r, o.y + q, o.w + s)] += 0.25 * g;
    }
}

// Adjoint of the exchange term over its source region
// V[o0 + p, o3 + s, o1 + q, o2 + r]: dV -= 1/4 sum_(a, b) G[(a,p), (b,q), (b,r), (a,s)].
__global__ void k_spin_2e_adj_exchange(const double* __restrict__ G, int64_t n, int4 o, int4 d,
                                       double* __restrict__ dV) {
    const int64_t D1 = 2 * d.y, D2 = 2 * d.z, D3 = 2 * d.w;
    XR_GRID_STRIDE(e, (int64_t)d.x * d.w * d.y * d.z) {
        int64_t t = e;                      // region order (p, s, q, r)
        const int r = (int)(t % d.z); t /= d.z;
        const int q = (int)(t % d.y); t /= d.y;
        const int s = (int)(t % d.w);
        const int p = (int)(t / d.w);
        double g = 0.0;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                g += G[((int64_t)(a * d.x + p) * D1 + b * d.y + q) * D2 * D3 + (int64_t)(b * d.z + r) * D3 + a * d.w + s];
        dV[v4(n, o.x + p, o.w + s, o.y + q, o.z + r)] -= 0.25 * g;
    }
}

// kinds[i]: 1 = one-body block (dims / offs entries 0 and 1 used), 2 =
// two-body block (entries 0..3); src[i] the spatial matrix or tensor
// (n x n or n^4), dst[i] the block (row-major).
extern "C" int xr_spin_blocks(xr_ctx* c, int n_keys, const int* kinds, const int* dims, const int* offs,
                              const double* const* src, double* const* dst, int64_t n) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n_keys < 0 || n <= 0 || (n_keys > 0 && (!kinds || !dims || !offs || !src || !dst)))
        return xr_fail(c, XR_EINVAL, "bad spin block list");
    ProfScope prof(c, XR_K_OTHER, 0.0, 0.0);
    for (int i = 0; i < n_keys; i++) {
        const int* d = dims + 4 * i;
        const int* o = offs + 4 * i;
        const int nd = kinds[i] == 1 ? 2 : 4;
        if (kinds[i] != 1 && kinds[i] != 2) return xr_fail(c, XR_EINVAL, "spin block kind must be 1 or 2");
        for (int a = 0; a < nd; a++)
            if (d[a] <= 0 || o[a] < 0 || o[a] + d[a] > n) return xr_fail(c, XR_ESHAPE, "spin block outside the integrals");
        if (kinds[i] == 1) {
            k_spin_1e<<<xr_grid(4LL * d[0] * d[1]), 256, 0, c->stream>>>(src[i], n, o[0], o[1], d[0], d[1], dst[i]);
        } else {
            const int4 oo = make_int4(o[0], o[1], o[2], o[3]), dd = make_int4(d[0], d[1], d[2], d[3]);
            k_spin_2e<<<xr_grid(16LL * d[0] * d[1] * d[2] * d[3]), 256, 0, c->stream>>>(src[i], n, oo, dd, dst[i]);
        }
        XR_LAUNCH_CHECK(c);
    }
    return XR_OK;
}

// Adds the adjoint of every block (grad[i], null: no gradient) into the
// spatial gradients dsrc[i] (blocks in list order, which fixes the sums).
extern "C" int xr_spin_blocks_adjoint(xr_ctx* c, int n_keys, const int* kinds, const int* dims, const int* offs,
                                      const double* const* grad, double* const* dsrc, int64_t n) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (n_keys < 0 || n <= 0 || (n_keys > 0 && (!kinds || !dims || !offs || !grad || !dsrc)))
        return xr_fail(c, XR_EINVAL, "bad spin block list");
    ProfScope prof(c, XR_K_OTHER, 0.0, 0.0);
    for (int i = 0; i < n_keys; i++) {
        if (!grad[i]) continue;
        const int* d = dims + 4 * i;
        const int* o = offs + 4 * i;
        const int nd = kinds[i] == 1 ? 2 : 4;
        if (kinds[i] != 1 && kinds[i] != 2) return xr_fail(c, XR_EINVAL, "spin block kind must be 1 or 2");
        for (int a = 0; a < nd; a++)
            if (d[a] <= 0 || o[a] < 0 || o[a] + d[a] > n) return xr_fail(c, XR_ESHAPE, "spin block outside the integrals");
        if (kinds[i] == 1) {
            k_spin_1e_adj<<<xr_grid((int64_t)d[0] * d[1]), 256, 0, c->stream>>>(grad[i], n, o[0], o[1], d[0], d[1], dsrc[i]);
            XR_LAUNCH_CHECK(c);
        } else {
            const int4 oo = make_int4(o[0], o[1], o[2], o[3]), dd = make_int4(d[0], d[1], d[2], d[3]);
            const int64_t cnt = (int64_t)d[0] * d[1] * d[2] * d[3];
            k_spin_2e_adj_direct<<<xr_grid(cnt), 256, 0, c->stream>>>(grad[i], n, oo, dd, dsrc[i]);
            XR_LAUNCH_CHECK(c);
            k_spin_2e_adj_exchange<<<xr_grid(cnt), 256, 0, c->stream>>>(grad[i], n, oo, dd, dsrc[i]);
            XR_LAUNCH_CHECK(c);
        }
    }
    return XR_OK;
}

// dst = src.permute(perm) for a row-major 4-index src of dims d (dst axis a
// is src axis perm[a]; dst row-major), written or added (accumulate): one
// thread per dst element (coalesced stores, gathered loads).
__global__ void k_permute4(const double* __restrict__ src, int4 d, int4 p, double* __restrict__ dst, int accumulate) {
    const int dd[4] = {d.x, d.y, d.z, d.w};
    const int pp[4] = {p.x, p.y, p.z, p.w};
    int64_t sst[4];                 // src strides
    sst[3] = 1;
    sst[2] = dd[3];
    sst[1] = (int64_t)dd[2] * dd[3];
    sst[0] = (int64_t)dd[1] * dd[2] * dd[3];
    const int e0 = dd[pp[0]], e1 = dd[pp[1]], e2 = dd[pp[2]], e3 = dd[pp[3]];
    const int64_t t0 = sst[pp[0]], t1 = sst[pp[1]], t2 = sst[pp[2]], t3 = sst[pp[3]];
    XR_GRID_STRIDE(e, (int64_t)e0 * e1 * e2 * e3) {
        int64_t t = e;
        const int64_t j3 = t % e3; t /= e3;
        const int64_t j2 = t % e2; t /= e2;
        const int64_t j1 = t % e1;
        const int64_t j0 = t / e1;
        const double v = src[j0 * t0 + j1 * t1 + j2 * t2 + j3 * t3];
        dst[e] = accumulate ? dst[e] + v : v;
    }
}

extern "C" int xr_permute4(xr_ctx* c, const double* src, const int64_t* dims, const int* perm, double* dst,
                           int accumulate) {
    if (!c) return XR_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!dims || !perm) return xr_fail(c, XR_EINVAL, "null dims / perm");
    int seen = 0;
    for (int a = 0; a < 4; a++) {
        if (dims[a] < 0 || dims[a] > 0x7fffffff) return xr_fail(c, XR_EINVAL, "bad dimension");
        if (perm[a] < 0 || perm[a] > 3 || (seen >> perm[a]) & 1) return xr_fail(c, XR_EINVAL, "perm is not a permutation of 0..3");
        seen |= 1 << perm[a];
    }
    const int64_t n = dims[0] * dims[1] * dims[2] * dims[3];
    if (n == 0) return XR_OK;
    if (src == dst) return xr_fail(c, XR_EINVAL, "xr_permute4 cannot work in place");
    ProfScope prof(c, XR_K_OTHER, 0.0, (accumulate ? 24.0 : 16.0) * n);
    k_permute4<<<xr_grid(n), 256, 0, c->stream>>>(src, make_int4((int)dims[0], (int)dims[1], (int)dims[2], (int)dims[3]),
                                                  make_int4(perm[0], perm[1], perm[2], perm[3]), dst, accumulate);
    XR_LAUNCH_CHECK(c);
    return XR_OK;
}
