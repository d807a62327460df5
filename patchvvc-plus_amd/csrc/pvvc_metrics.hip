/*
 * pvvc_metrics.hip -- quality metric on the GPU: the symmetric
 * nearest-neighbour MSE of vvc::common::MSE::Compute
 * (src/registration/registration_base.cc:168-223) and the PSNR that
 * PVVCDecompression prints from it (src/codec/pvvc_decompression.cc:150-164).
 *
 * The reference builds two PCL k-d trees and runs one nearestKSearch per
 * point on the CPU.  Here each direction is:
 *
 *   1. the target cloud's bounding box (block min/max, one small read-back);
 *   2. a uniform grid of cubic cells of side h = 2 * extent / sqrt(n): a
 *      voxelised surface puts a few points in each non-empty cell;
 *   3. (cell key, index) pairs radix-sorted (rocPRIM), the points gathered
 *      into that order as float4 {x, y, z, index} so a cell is one coalesced
 *      run, and the head of every run entered into an open-addressing hash
 *      table (key -> start, count);
 *   4. one thread per query, queries taken in the sorted order of their own
 *      grid so a wavefront walks one region: shells of cells of Chebyshev
 *      radius 0, 1, 2, ... until no unvisited cell can hold a point at or
 *      under the best distance (the bound is the distance from the query to
 *      the faces of the searched cube, kept conservative for FP32 rounding);
 *      past kMaxShell shells a query scans every target point instead
 *      (a cloud degenerate to one cell scans it whole: O(n^2));
 *   5. per query the FP32 squared distance and the FP64 squares of the FP32
 *      Y/U/V differences, reduced per block in LDS and then by one block, in
 *      a fixed order: the sums are deterministic.
 *
 * The work is HBM/L2-latency bound gather work (a few cache lines per
 * query), not arithmetic.
 */
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "pvvc_runtime.h"

namespace pvvc {
namespace metric {

constexpr int kBlock = 256;
constexpr int kMaxShell = 8;
constexpr uint64_t kEmpty = ~0ull;

struct GridDev {
    double org[3];
    double h;
    int64_t dim[3];
    const uint64_t* skey;  /* sorted cell keys */
    const float4* spt;     /* points in sorted order; w = original index (bits) */
    const uint64_t* hkey;  /* hash table keys (kEmpty = free) */
    const uint2* hval;     /* (first position, count) of the key's run */
    uint32_t hmask;
    uint32_t n;
};

__host__ __device__ inline uint64_t pack_cell(int64_t x, int64_t y, int64_t z) {
    return (uint64_t(x) << 42) | (uint64_t(y) << 21) | uint64_t(z);
}

__device__ inline uint32_t hash_slot(uint64_t k, uint32_t mask) {
    return static_cast<uint32_t>((k * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

__device__ inline void cell_of(const GridDev& g, const float* p, int64_t* c) {
    for (int a = 0; a < 3; ++a) {
        const int64_t v = static_cast<int64_t>(floor((double(p[a]) - g.org[a]) / g.h));
        c[a] = v < 0 ? 0 : (v > g.dim[a] - 1 ? g.dim[a] - 1 : v);
    }
}

/* (dx*dx + dy*dy) + dz*dz in FP32 (the library builds with no contraction) */
__device__ inline float sq_dist(float ax, float ay, float az, float bx, float by, float bz) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    return (dx * dx + dy * dy) + dz * dz;
}

/* ColorYUV(const pcl::PointXYZRGB&), include/common/common.h:95-99 */
__device__ inline void yuv_of(const uint8_t* c, float* o) {
    const float r = c[0], g = c[1], b = c[2];
    o[0] = r * 0.299f + g * 0.587f + b * 0.114f;
    o[1] = r * -0.168736f - g * 0.331264f + b * 0.5f + 128.0f;
    o[2] = r * 0.5f - g * 0.418688f - b * 0.081312f + 128.0f;
}

/* per-block min/max of x, y, z: out[block] = {min xyz, max xyz} */
__global__ void bbox_kernel(const float* __restrict__ xyz, uint32_t n, float* __restrict__ out) {
    __shared__ float red[6][kBlock];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        for (int a = 0; a < 3; ++a) {
            const float v = xyz[3ull * i + a];
            lo[a] = fminf(lo[a], v);
            hi[a] = fmaxf(hi[a], v);
        }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6ull * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void key_kernel(GridDev g, const float* __restrict__ xyz, uint32_t n, uint64_t* __restrict__ key,
                           uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float p[3] = {xyz[3ull * i], xyz[3ull * i + 1], xyz[3ull * i + 2]};
    int64_t c[3];
    cell_of(g, p, c);
    key[i] = pack_cell(c[0], c[1], c[2]);
    idx[i] = i;
}

/* points into sorted order; run heads into the hash table */
__global__ void insert_kernel(const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                              const float* __restrict__ xyz, uint32_t n, float4* __restrict__ spt,
                              uint64_t* hkey, uint2* __restrict__ hval, uint32_t mask) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t src = sidx[i];
    spt[i] = make_float4(xyz[3ull * src], xyz[3ull * src + 1], xyz[3ull * src + 2], __uint_as_float(src));
    const uint64_t k = skey[i];
    if (i > 0 && skey[i - 1] == k) return;
    uint32_t e = i + 1;
    while (e < n && skey[e] == k) ++e;
    uint32_t s = hash_slot(k, mask);
    for (;;) { /* the table is at most half full, so a free slot exists */
        const unsigned long long prev =
            atomicCAS(reinterpret_cast<unsigned long long*>(hkey + s), static_cast<unsigned long long>(kEmpty),
                      static_cast<unsigned long long>(k));
        if (prev == kEmpty) {
            hval[s] = make_uint2(i, e - i);
            return;
        }
        s = (s + 1) & mask;
    }
}

__device__ inline void scan_run(const GridDev& g, uint2 run, float qx, float qy, float qz, float& best,
                                uint32_t& bi) {
    for (uint32_t j = run.x; j < run.x + run.y; ++j) {
        const float4 t = g.spt[j];
        const float d = sq_dist(qx, qy, qz, t.x, t.y, t.z);
        const uint32_t ti = __float_as_uint(t.w);
        if (d < best || (d == best && ti < bi)) {
            best = d;
            bi = ti;
        }
    }
}

__device__ inline bool lookup(const GridDev& g, uint64_t k, uint2& run) {
    uint32_t s = hash_slot(k, g.hmask);
    for (;;) {
        const uint64_t v = g.hkey[s];
        if (v == k) {
            run = g.hval[s];
            return true;
        }
        if (v == kEmpty) return false;
        s = (s + 1) & g.hmask;
    }
}

/* One direction of MSE::Compute.  Query i is qorder[i] (its own grid's sorted
 * order).  partial[block] = {geo, y, u, v} sums of the block's queries. */
__global__ void __launch_bounds__(kBlock) query_kernel(GridDev g, const float* __restrict__ qxyz,
                                                       const uint8_t* __restrict__ qrgb,
                                                       const uint32_t* __restrict__ qorder, uint32_t nq,
                                                       const uint8_t* __restrict__ trgb, int64_t* __restrict__ nn,
                                                       double* __restrict__ partial) {
    __shared__ double red[4][kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double term[4] = {0, 0, 0, 0};
    if (i < nq) {
        const uint32_t qi = qorder[i];
        const float q[3] = {qxyz[3ull * qi], qxyz[3ull * qi + 1], qxyz[3ull * qi + 2]};
        int64_t c[3];
        cell_of(g, q, c);
        float best = INFINITY;
        uint32_t bi = 0xffffffffu;
        bool done = false;
        const double slack = 1e-9 * (fabs(double(q[0])) + fabs(double(q[1])) + fabs(double(q[2])) + g.h);
        for (int r = 0; r <= kMaxShell && !done; ++r) {
            for (int dx = -r; dx <= r; ++dx) {
                const int64_t cx = c[0] + dx;
                if (cx < 0 || cx >= g.dim[0]) continue;
                for (int dy = -r; dy <= r; ++dy) {
                    const int64_t cy = c[1] + dy;
                    if (cy < 0 || cy >= g.dim[1]) continue;
                    /* interior columns of the shell hold only their two end cells */
                    const bool face = (dx == -r || dx == r || dy == -r || dy == r);
                    const int step = face ? 1 : 2 * r;
                    for (int dz = -r; dz <= r; dz += step) {
                        const int64_t cz = c[2] + dz;
                        if (cz >= 0 && cz < g.dim[2]) {
                            uint2 run;
                            if (lookup(g, pack_cell(cx, cy, cz), run)) scan_run(g, run, q[0], q[1], q[2], best, bi);
                        }
                    }
                }
            }
            /* every unvisited point lies outside cells [c - r, c + r] */
            double lb = INFINITY;
            for (int a = 0; a < 3; ++a) {
                if (c[a] - r > 0) lb = fmin(lb, double(q[a]) - (g.org[a] + double(c[a] - r) * g.h));
                if (c[a] + r < g.dim[a] - 1) lb = fmin(lb, g.org[a] + double(c[a] + r + 1) * g.h - double(q[a]));
            }
            if (lb == INFINITY) {
                done = true; /* the searched cube covers the grid */
            } else {
                lb -= slack;
                if (lb > 0 && double(best) < lb * lb * (1.0 - 1e-6)) done = true;
            }
        }
        if (!done) /* far from every target point: scan them all */
            scan_run(g, make_uint2(0u, g.n), q[0], q[1], q[2], best, bi);
        float a[3], b[3];
        yuv_of(qrgb + 3ull * qi, a);
        yuv_of(trgb + 3ull * bi, b);
        term[0] = double(best);
        for (int k = 0; k < 3; ++k) {
            const double d = double(a[k] - b[k]);
            term[k + 1] = d * d;
        }
        if (nn) nn[qi] = bi;
    }
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = term[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[4ull * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

/* out[k] = sum of partial[b][k] over blocks, fixed order (one block) */
__global__ void reduce_kernel(const double* __restrict__ partial, uint32_t blocks, double* __restrict__ out) {
    __shared__ double red[4][kBlock];
    double s[4] = {0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock)
        for (int k = 0; k < 4; ++k) s[k] += partial[4ull * b + k];
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

/* Device allocations of one call, freed on every path. */
struct Allocs {
    std::vector<void*> p;
    ~Allocs() {
        for (void* x : p) (void)hipFree(x);
    }
    template <class T>
    T* get(size_t count) {
        void* d = nullptr;
        if (!rt_hip_ok(hipMalloc(&d, std::max<size_t>(count, 1) * sizeof(T)))) return nullptr;
        p.push_back(d);
        return static_cast<T*>(d);
    }
};

/* A built grid over one cloud. */
struct Grid {
    GridDev dev{};
    uint32_t* order = nullptr; /* original indices in sorted order */
};

int build_grid(const float* xyz, uint32_t n, Allocs& al, hipStream_t s, Grid& out) {
    const uint32_t nb = std::min<uint32_t>((n + kBlock - 1) / kBlock, 1024);
    float* bb = al.get<float>(6ull * nb);
    if (!bb) return PVVC_ERR_HIP;
    hipLaunchKernelGGL(bbox_kernel, dim3(nb), dim3(kBlock), 0, s, xyz, n, bb);
    std::vector<float> hb(6ull * nb);
    if (!rt_hip_ok(hipGetLastError()) ||
        !rt_hip_ok(hipMemcpyAsync(hb.data(), bb, hb.size() * sizeof(float), hipMemcpyDeviceToHost, s)) ||
        !rt_hip_ok(hipStreamSynchronize(s)))
        return PVVC_ERR_HIP;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], double(hb[6 * b + a]));
            hi[a] = std::max(hi[a], double(hb[6 * b + 3 + a]));
        }
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(lo[a]) || !std::isfinite(hi[a])) return PVVC_ERR_INVALID_ARG; /* NaN / inf input */
    GridDev& g = out.dev;
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6});
    g.h = std::max(2.0 * ext / std::sqrt(double(n)), ext / 1.0e6);
    for (int a = 0; a < 3; ++a) {
        g.org[a] = lo[a];
        g.dim[a] = std::min<int64_t>(static_cast<int64_t>((hi[a] - lo[a]) / g.h) + 1, (int64_t(1) << 21) - 1);
    }
    g.n = n;
    uint64_t* key = al.get<uint64_t>(n);
    uint64_t* skey = al.get<uint64_t>(n);
    uint32_t* idx = al.get<uint32_t>(n);
    uint32_t* sidx = al.get<uint32_t>(n);
    float4* spt = al.get<float4>(n);
    uint32_t hsize = 1024;
    while (hsize < 2ull * n) hsize <<= 1;
    uint64_t* hkey = al.get<uint64_t>(hsize);
    uint2* hval = al.get<uint2>(hsize);
    if (!key || !skey || !idx || !sidx || !spt || !hkey || !hval) return PVVC_ERR_HIP;
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(key_kernel, dim3(grid), dim3(kBlock), 0, s, g, xyz, n, key, idx);
    size_t tmp_bytes = 0;
    if (!rt_hip_ok(rocprim::radix_sort_pairs(nullptr, tmp_bytes, key, skey, idx, sidx, n, 0u, 63u, s)))
        return PVVC_ERR_HIP;
    void* tmp = al.get<uint8_t>(tmp_bytes);
    if (!tmp || !rt_hip_ok(rocprim::radix_sort_pairs(tmp, tmp_bytes, key, skey, idx, sidx, n, 0u, 63u, s)))
        return PVVC_ERR_HIP;
    if (!rt_hip_ok(hipMemsetAsync(hkey, 0xff, sizeof(uint64_t) * hsize, s))) return PVVC_ERR_HIP;
    hipLaunchKernelGGL(insert_kernel, dim3(grid), dim3(kBlock), 0, s, skey, sidx, xyz, n, spt, hkey, hval,
                       hsize - 1);
    if (!rt_hip_ok(hipGetLastError())) return PVVC_ERR_HIP;
    g.skey = skey;
    g.spt = spt;
    g.hkey = hkey;
    g.hval = hval;
    g.hmask = hsize - 1;
    out.order = sidx;
    return PVVC_OK;
}

int direction(const Grid& tgt, const Grid& qry, const float* qxyz, const uint8_t* qrgb, uint32_t nq,
              const uint8_t* trgb, int64_t* nn, double* sums_dev, double* partial, hipStream_t s) {
    const uint32_t blocks = (nq + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(query_kernel, dim3(blocks), dim3(kBlock), 0, s, tgt.dev, qxyz, qrgb, qry.order, nq, trgb, nn,
                       partial);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kBlock), 0, s, partial, blocks, sums_dev);
    return rt_hip_ok(hipGetLastError()) ? PVVC_OK : PVVC_ERR_HIP;
}

}  // namespace metric
}  // namespace pvvc

using namespace pvvc;

extern "C" {

int pvvc_cloud_mse(const float* p_xyz, const uint8_t* p_rgb, uint64_t np, const float* q_xyz, const uint8_t* q_rgb,
                   uint64_t nq, int64_t* nn_p, int64_t* nn_q, pvvc_mse* out, double* kernel_ms) {
    if (!out) return PVVC_ERR_INVALID_ARG;
    if (kernel_ms) *kernel_ms = 0.0;
    /* MSE::Compute (registration_base.cc:170-179): an empty cloud gives -1 for
     * its direction and the computation stops with the other one at 0. */
    for (int k = 0; k < 4; ++k) out->p_to_q[k] = out->q_to_p[k] = 0.0;
    if (np == 0)
        for (int k = 0; k < 4; ++k) out->p_to_q[k] = -1.0;
    if (nq == 0)
        for (int k = 0; k < 4; ++k) out->q_to_p[k] = -1.0;
    if (np == 0 || nq == 0) return PVVC_OK;
    if (!p_xyz || !p_rgb || !q_xyz || !q_rgb || np >= (1ull << 31) || nq >= (1ull << 31))
        return PVVC_ERR_INVALID_ARG;
    hipStream_t s = rt_stream();
    metric::Allocs al;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rt_hip_ok(hipEventCreate(&e0)) || !rt_hip_ok(hipEventCreate(&e1))) {
        if (e0) (void)hipEventDestroy(e0);
        return PVVC_ERR_HIP;
    }
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    if (!rt_hip_ok(hipEventRecord(e0, s))) return PVVC_ERR_HIP;
    metric::Grid gp, gq;
    int rc = metric::build_grid(p_xyz, static_cast<uint32_t>(np), al, s, gp);
    if (rc == PVVC_OK) rc = metric::build_grid(q_xyz, static_cast<uint32_t>(nq), al, s, gq);
    if (rc != PVVC_OK) return rc;
    const uint32_t bmax = static_cast<uint32_t>((std::max(np, nq) + metric::kBlock - 1) / metric::kBlock);
    double* partial = al.get<double>(4ull * bmax);
    double* sums = al.get<double>(8);
    if (!partial || !sums) return PVVC_ERR_HIP;
    rc = metric::direction(gq, gp, p_xyz, p_rgb, static_cast<uint32_t>(np), q_rgb, nn_p, sums, partial, s);
    if (rc == PVVC_OK)
        rc = metric::direction(gp, gq, q_xyz, q_rgb, static_cast<uint32_t>(nq), p_rgb, nn_q, sums + 4, partial, s);
    if (rc != PVVC_OK) return rc;
    double h[8];
    if (!rt_hip_ok(hipEventRecord(e1, s)) ||
        !rt_hip_ok(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s)) ||
        !rt_hip_ok(hipStreamSynchronize(s)))
        return PVVC_ERR_HIP;
    if (kernel_ms) {
        /* from the first launch to the last, including the two bounding-box
         * read-backs in between */
        float ms = 0.0f;
        if (rt_hip_ok(hipEventElapsedTime(&ms, e0, e1))) *kernel_ms = ms;
    }
    for (int k = 0; k < 4; ++k) {
        out->p_to_q[k] = h[k] / double(np);
        out->q_to_p[k] = h[4 + k] / double(nq);
    }
    return PVVC_OK;
}

float pvvc_psnr(double mse_a, double mse_b, float peak) {
    const float m = static_cast<float>(std::max(mse_a, mse_b));
    return 10.0f * std::log10(peak * peak / m);
}

}  // extern "C"
