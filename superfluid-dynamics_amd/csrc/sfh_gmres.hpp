// sfh_gmres.hpp -- GMRES state shared by the Arnoldi kernel and the kernels that do GMRES work on
// the side (the assembly kernel forms the initial residual while it writes the matrix; the
// derivative kernel re-arms the control block and writes the extrapolated initial guess).
#pragma once
#include <cstddef>

#include "sfh_kernels.hpp"

namespace sfh {

constexpr int kRowsPerWG = 4;      // one matrix row per wave
constexpr int kMaxKrylov = 32;     // basis vectors per solve (iteration cap)
constexpr int kGmMaxN = 4096;      // LDS copy of the vector being multiplied

struct GmCtl {
    int done;          // 0 while the solve runs, else 1 + the launch that finished it
    int refine_iters;  // Arnoldi steps of the refinement solve (mixed precision; 0 when none ran)
    int iters;         // Arnoldi steps of the last solve
    int unconverged;   // solves that ended at the launch limit (cumulative)
    double bnorm, beta, res;
    double hnext;
    int fill_iters;    // most Arnoldi steps of a solve since the stage history last held fewer
                       // than kHistDepth steps
    int spare[19];
};
static_assert(sizeof(GmCtl) == kGmCtlInts * sizeof(int), "host reads control blocks with this stride");

struct GmView {
    const double* M;
    const float* M32;
    const double* b;
    const double* x;   // initial guess
    double* sol;       // solution
    double* V;       // (m+1) x n, unnormalised: v_i = sc[i] * V[i]
    double* w0;      // products of even steps
    double* w1;      // products of odd steps
    double* sc;      // m+1 scales
    double* gc;      // m+1: component i of the rotated right-hand side before rotation i
    double* cs;
    double* sn;
    double* g;
    double* H;       // column j at H[j*(m+1)]
    double* y;       // m+1: y_q sc_q of a solve that converged before the last launch of its sequence
    double* part;    // 2 slots (launch parity) of (m+1) x nwg partial sums, basis index major
    GmCtl* ctl;
    int n, nwg;
};

__device__ inline GmView gm_view(const GmArgs& a, const SurfDesc* sd, int s) {
    const SurfDesc d = sd[s];
    GmView v;
    const int m = a.m;
    v.n = d.n;
    v.nwg = (d.n + kRowsPerWG - 1) / kRowsPerWG;
    v.M = a.M + d.moff;
    v.M32 = a.M32 ? a.M32 + d.moff : nullptr;
    v.b = a.b + d.voff;
    v.x = a.x + d.voff;
    v.sol = a.sol + d.voff;
    v.V = a.V + (long long)(m + 1) * d.voff;
    v.w0 = a.w + d.voff;
    v.w1 = a.w + a.total + d.voff;
    v.sc = a.small + (long long)s * a.small_stride;
    v.gc = v.sc + (m + 1);
    v.cs = v.gc + (m + 1);
    v.sn = v.cs + m;
    v.g = v.sn + m;
    v.H = v.g + (m + 1);
    v.y = v.H + (long long)m * (m + 1);
    v.part = a.part + (long long)s * a.part_stride;
    v.ctl = reinterpret_cast<GmCtl*>(a.ctl) + s;
    return v;
}

__device__ inline GmView gm_view(const GmArgs& a, const SurfDesc* sd) { return gm_view(a, sd, blockIdx.z); }

// Arnoldi steps of the solve, in the field of its kind (a refinement solve keeps the first solve's)
__device__ inline void set_iters(GmCtl* ctl, const GmArgs& a, int k) {
    ctl->fill_iters = max(ctl->fill_iters, k);
    if (a.check_first)
        ctl->refine_iters = k;
    else
        ctl->iters = k;
}

// This surface's history of solutions of the current RK stage (null for a plain evaluation), and
// the stage's count of stored steps (one increment per solve, by surface 0).
__device__ inline double* history(const GmArgs& a, const SurfDesc* sd, int s) {
    return a.stage > 0 ? a.hist + (long long)(a.stage - 1) * kHistDepth * a.total + sd[s].voff : nullptr;
}
__device__ inline void count_history(const GmArgs& a) {
    a.hvalid[a.stage - 1] = min(a.hvalid[a.stage - 1] + 1, kHistDepth);
}

// Sums part[i * nwg + w] over w for i < count (count <= kMaxKrylov + 1) into out[i] (LDS).
// Basis indices go in chunks of 12; within a chunk every load of a thread (12 indices x 2
// workgroups per 512-workgroup pass) is issued before any addition.
__device__ inline void reduce_partials(const double* __restrict__ part, int nwg, int count, double* out) {
    constexpr int CH = 12;
    __shared__ double red[4][kMaxKrylov + 2];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i0 = 0; i0 < count; i0 += CH) {
        double acc[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) acc[u] = 0.0;
        for (int base = 0; base < nwg; base += 512) {
            const int w0 = base + t, w1 = base + t + 256;
            double v0[CH], v1[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const bool ok = i0 + u < count;
                v0[u] = (ok && w0 < nwg) ? part[(i0 + u) * nwg + w0] : 0.0;
                v1[u] = (ok && w1 < nwg) ? part[(i0 + u) * nwg + w1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) acc[u] += v0[u] + v1[u];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (i0 + u < count) {
                const double s = wave_sum(acc[u]);
                if (lane == 0) red[wv][i0 + u] = s;
            }
        }
    }
    __syncthreads();
    if (t < count) out[t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    __syncthreads();
}

// The same sums for count <= kFastDots and nwg <= 512, split in two so that a kernel can issue
// the loads together with its other early loads and add them up when it needs the sums (a wait
// for one load waits for every load issued before it).
constexpr int kFastDots = 4;
struct DotLoads {
    double a[kFastDots], b[kFastDots];
};
__device__ inline bool dots_fast(int nwg, int count) { return count <= kFastDots && nwg <= 512; }
__device__ inline void dots_load(const double* __restrict__ part, int nwg, int count, DotLoads& L) {
    const int w0 = threadIdx.x, w1 = threadIdx.x + 256;
#pragma unroll
    for (int u = 0; u < kFastDots; ++u) {
        const bool ok = u < count;
        L.a[u] = (ok && w0 < nwg) ? part[u * nwg + w0] : 0.0;
        L.b[u] = (ok && w1 < nwg) ? part[u * nwg + w1] : 0.0;
    }
}
__device__ inline void dots_finish(const DotLoads& L, int count, double* out) {
    __shared__ double red[4][kFastDots];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int u = 0; u < kFastDots; ++u) {
        if (u < count) {
            const double s = wave_sum(L.a[u] + L.b[u]);
            if (lane == 0) red[wv][u] = s;
        }
    }
    __syncthreads();
    if (t < count) out[t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    __syncthreads();
}

__device__ inline double block_sum4(double v, double* red4) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) red4[wv] = v;
    __syncthreads();
    const double s = red4[0] + red4[1] + red4[2] + red4[3];
    __syncthreads();
    return s;
}

}  // namespace sfh
