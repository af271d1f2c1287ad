// gmres.hip -- FP64 GMRES for M mu = b over the stored dipole matrix (the north-star alternative
// to the LU of MatrixSolver::solve, CuSuperHelium/MatrixSolver.cuh:113-172).
//
// M = I/2 + K is a second-kind operator: its spectrum clusters at 1/2 (for the helium film one
// outlier sits far to the right), so GMRES converges in a handful of iterations, and the
// stepper warm starts every solve from the previous stage's mu.  Each iteration is two kernels
// (classical Gram-Schmidt, one pass):
//   gm_arnoldi : w = M v_j and the dots h_i = v_i . w (i <= j)
//   gm_extend  : u = w - sum h_i v_i (stored unnormalised, scale 1/|u| kept aside) and |u|
//                -> last workgroup: h_{j+1,j} = |u|, Givens rotation, residual test
// The new vector's norm is reduced from u itself (not from |w|^2 - sum h_i^2, which cancels
// catastrophically once the basis drifts from orthonormal).  One pass suffices here: the basis
// conditioning follows that of M = I/2 + K, which is O(1) for the configurations this solves.
// Reductions use the "last workgroup done" pattern (atomic counter after a release fence), so
// no workgroup ever waits for another.  Once the residual estimate reaches tol * |b| the
// remaining launches of the fixed-length sequence return at once.
//
// Layout for memory-level parallelism: a workgroup owns 8 rows; its 256 threads are (row r,
// basis index i) pairs (r = 2 * wave + lane / 32, i = lane % 32), so every basis-vector access of
// an iteration is one load per thread, all in flight together.  Partial sums are stored basis
// index major (part[i * nwg + wg]) and the last workgroup reduces all of them with unrolled loads.
#include "sfh_kernels.hpp"

namespace sfh {

constexpr int kRowsPerWG = 8;    // 4 waves x 2 rows
constexpr int kMaxKrylov = 32;   // basis vectors per solve (iteration cap)

struct GmCtl {
    int done;
    int counter;
    int iters;
    int unconverged;
    double bnorm, beta, res;
    double hnext;
    int pad[4];
};
static_assert(sizeof(GmCtl) == kGmCtlInts * sizeof(int), "host reads control blocks with this stride");

struct GmView {
    const double* M;
    const double* b;
    double* x;
    double* V;       // (m+1) x n, unnormalised
    double* sc;      // m+1 scales: v_i = sc[i] * V[i]
    double* w;       // n
    double* part;    // (m+2) x nwg
    double* H;       // column j at H[j*(m+1)]
    double* h1;      // m+1
    double* cs;
    double* sn;
    double* g;
    GmCtl* ctl;
    int n, nwg;
};

__device__ inline GmView gm_view(const GmArgs& a, const SurfDesc* sd) {
    const int s = blockIdx.z;
    const SurfDesc d = sd[s];
    GmView v;
    const int m = a.m;
    v.n = d.n;
    v.nwg = (d.n + kRowsPerWG - 1) / kRowsPerWG;
    v.M = a.M + d.moff;
    v.b = a.b + d.voff;
    v.x = a.x + d.voff;
    v.V = a.V + (long long)(m + 1) * d.voff;
    v.w = a.w + d.voff;
    v.sc = a.small + (long long)s * a.small_stride;
    v.h1 = v.sc + (m + 1);
    v.cs = v.h1 + (m + 1);
    v.sn = v.cs + m;
    v.g = v.sn + m;
    v.H = v.g + (m + 1);
    v.part = a.part + (long long)s * a.part_stride;
    v.ctl = reinterpret_cast<GmCtl*>(a.ctl) + s;
    return v;
}

// Dot products of two matrix rows with one vector, one wave, 8 loads per row in flight.
__device__ inline void rows_dot2(const double* __restrict__ R0, const double* __restrict__ R1,
                                 const double* __restrict__ vec, int n, int lane, double& d0, double& d1) {
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
    int c = lane;
    for (; c + 7 * 64 < n; c += 8 * 64) {
        double m0[8], m1[8], vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            m0[u] = R0[c + u * 64];
            m1[u] = R1[c + u * 64];
            vv[u] = vec[c + u * 64];
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            a0 += m0[u] * vv[u];
            a1 += m0[u + 1] * vv[u + 1];
            b0 += m1[u] * vv[u];
            b1 += m1[u + 1] * vv[u + 1];
        }
    }
    for (; c < n; c += 64) {
        a0 += R0[c] * vec[c];
        b0 += R1[c] * vec[c];
    }
    d0 = wave_sum(a0 + a1);
    d1 = wave_sum(b0 + b1);
}

// Returns true on the workgroup that finished last (the counter is then re-armed).  Every
// workgroup increments with release semantics (its partial sums become visible at agent scope);
// only the last one pays for the acquire.  On gfx950 an agent-scope acquire invalidates the L2 of
// the executing XCD, so a full fence in every workgroup would evict the data other workgroups are
// streaming.
__device__ inline bool last_workgroup(GmCtl* ctl, int nwg) {
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int prev = __hip_atomic_fetch_add(&ctl->counter, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == nwg - 1);
        if (is_last) {
            __hip_atomic_store(&ctl->counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
    return is_last != 0;
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

// r0 = b - M x0 into V[0]; last workgroup sets beta = |r0|, |b|, g[0], done.
__global__ __launch_bounds__(256) void gm_init_kernel(GmArgs a, const SurfDesc* sd) {
    const GmView v = gm_view(a, sd);
    if ((int)blockIdx.x >= v.nwg) return;
    __shared__ double rr[kRowsPerWG], bb[kRowsPerWG];
    __shared__ double sums[kMaxKrylov + 2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, n = v.n;
    const int r0 = blockIdx.x * kRowsPerWG + wv * 2;
    const int ra = min(r0, n - 1), rb = min(r0 + 1, n - 1);
    double d0, d1;
    rows_dot2(v.M + (long long)ra * n, v.M + (long long)rb * n, v.x, n, lane, d0, d1);
    if (lane == 0) {
        for (int q = 0; q < 2; ++q) {
            const int row = r0 + q;
            double r = 0.0, bv = 0.0;
            if (row < n) {
                bv = v.b[row];
                r = bv - (q ? d1 : d0);
                v.V[row] = r;
            }
            rr[wv * 2 + q] = r * r;
            bb[wv * 2 + q] = bv * bv;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < kRowsPerWG; ++i) {
            s1 += rr[i];
            s2 += bb[i];
        }
        v.part[blockIdx.x] = s1;
        v.part[v.nwg + blockIdx.x] = s2;
    }
    if (!last_workgroup(v.ctl, v.nwg)) return;
    reduce_partials(v.part, v.nwg, 2, sums);
    if (threadIdx.x == 0) {
        const double beta = sqrt(sums[0]), bnorm = sqrt(sums[1]);
        GmCtl* c = v.ctl;
        c->beta = beta;
        c->bnorm = bnorm;
        c->res = beta;
        c->iters = 0;
        c->done = (beta <= a.tol * bnorm) ? 1 : 0;
        v.g[0] = beta;
        v.sc[0] = beta > 0.0 ? 1.0 / beta : 0.0;
    }
}

// Arnoldi step j, first half: w = M v_j; h_i = v_i . w for i <= j (into h1 and column j of H).
__global__ __launch_bounds__(256) void gm_arnoldi_kernel(GmArgs a, const SurfDesc* sd, int j) {
    const GmView v = gm_view(a, sd);
    if ((int)blockIdx.x >= v.nwg || v.ctl->done) return;
    __shared__ double wl[kRowsPerWG];
    __shared__ double scratch[kRowsPerWG * 32];
    __shared__ double sums[kMaxKrylov + 2];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, n = v.n, m = a.m;
    const int row0 = blockIdx.x * kRowsPerWG;
    const int r = wv * 2 + (lane >> 5), i = lane & 31;
    const int row = row0 + r;
    // this thread's basis element, fetched before the long matvec
    const double vi = (i <= j && row < n) ? v.V[(long long)i * n + row] : 0.0;
    const double sj = v.sc[j];
    {
        const int ra = min(row0 + wv * 2, n - 1), rb = min(row0 + wv * 2 + 1, n - 1);
        double d0, d1;
        rows_dot2(v.M + (long long)ra * n, v.M + (long long)rb * n, v.V + (long long)j * n, n, lane, d0, d1);
        if (lane == 0) {
            wl[wv * 2] = (row0 + wv * 2 < n) ? sj * d0 : 0.0;
            wl[wv * 2 + 1] = (row0 + wv * 2 + 1 < n) ? sj * d1 : 0.0;
        }
    }
    __syncthreads();
    if (t < kRowsPerWG && row0 + t < n) v.w[row0 + t] = wl[t];
    scratch[r * 32 + i] = (i <= j) ? vi * wl[r] : 0.0;
    __syncthreads();
    if (t <= j) {
        double s = 0.0;
#pragma unroll
        for (int rr = 0; rr < kRowsPerWG; ++rr) s += scratch[rr * 32 + t];
        v.part[t * v.nwg + blockIdx.x] = s * v.sc[t];
    }
    if (!last_workgroup(v.ctl, v.nwg)) return;
    reduce_partials(v.part, v.nwg, j + 1, sums);
    if (t <= j) {
        v.h1[t] = sums[t];
        v.H[(long long)j * (m + 1) + t] = sums[t];
    }
}

// Arnoldi step j, second half: u = w - sum_{i<=j} h_i v_i into V[j+1]; the last workgroup takes
// h_{j+1,j} = |u|, applies the Givens rotations to column j and tests the residual.
__global__ __launch_bounds__(256) void gm_extend_kernel(GmArgs a, const SurfDesc* sd, int j) {
    const GmView v = gm_view(a, sd);
    if ((int)blockIdx.x >= v.nwg || v.ctl->done) return;
    __shared__ double ul[kRowsPerWG];
    __shared__ double sums[kMaxKrylov + 2];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, n = v.n, m = a.m;
    const int row0 = blockIdx.x * kRowsPerWG;
    const int r = wv * 2 + (lane >> 5), i = lane & 31;
    const int row = row0 + r;
    const double vi = (i <= j && row < n) ? v.V[(long long)i * n + row] : 0.0;
    const double wr = row < n ? v.w[row] : 0.0;
    double p = (i <= j) ? v.h1[i] * v.sc[i] * vi : 0.0;
    for (int o = 16; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
    const double u = row < n ? wr - p : 0.0;
    if (i == 0) {
        if (row < n) v.V[(long long)(j + 1) * n + row] = u;
        ul[r] = u * u;
    }
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
#pragma unroll
        for (int rr = 0; rr < kRowsPerWG; ++rr) s += ul[rr];
        v.part[blockIdx.x] = s;
    }
    if (!last_workgroup(v.ctl, v.nwg)) return;
    reduce_partials(v.part, v.nwg, 1, sums);
    if (t == 0) {
        GmCtl* c = v.ctl;
        double* Hj = v.H + (long long)j * (m + 1);
        const double hn = sqrt(sums[0]);
        Hj[j + 1] = hn;
        for (int q = 0; q < j; ++q) {
            const double t0 = v.cs[q] * Hj[q] + v.sn[q] * Hj[q + 1];
            const double t1 = -v.sn[q] * Hj[q] + v.cs[q] * Hj[q + 1];
            Hj[q] = t0;
            Hj[q + 1] = t1;
        }
        const double den = hypot(Hj[j], Hj[j + 1]);
        const double cj = den > 0.0 ? Hj[j] / den : 1.0;
        const double sn = den > 0.0 ? Hj[j + 1] / den : 0.0;
        v.cs[j] = cj;
        v.sn[j] = sn;
        Hj[j] = den;
        Hj[j + 1] = 0.0;
        v.g[j + 1] = -sn * v.g[j];
        v.g[j] = cj * v.g[j];
        c->res = fabs(v.g[j + 1]);
        c->iters = j + 1;
        c->hnext = hn;
        v.sc[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
        if (c->res <= a.tol * c->bnorm || hn == 0.0) {
            c->done = 1;
        } else if (j + 1 >= m) {
            c->done = 1;
            c->unconverged += 1;
        }
    }
}

// x += sum_i y_i v_i with R y = g (R = rotated H) solved by every workgroup redundantly.
__global__ __launch_bounds__(256) void gm_finish_kernel(GmArgs a, const SurfDesc* sd) {
    const GmView v = gm_view(a, sd);
    const int n = v.n, m = a.m;
    const int k = v.ctl->iters;
    if (k == 0) return;
    __shared__ double y[kMaxKrylov];
    __shared__ double ys[kMaxKrylov];
    if (threadIdx.x == 0) {
        for (int i = k - 1; i >= 0; --i) {
            double s = v.g[i];
            for (int l = i + 1; l < k; ++l) s -= v.H[(long long)l * (m + 1) + i] * y[l];
            y[i] = s / v.H[(long long)i * (m + 1) + i];
        }
        for (int i = 0; i < k; ++i) ys[i] = y[i] * v.sc[i];
    }
    __syncthreads();
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
        double vv[kMaxKrylov];
#pragma unroll
        for (int i = 0; i < kMaxKrylov; ++i) vv[i] = (i < k) ? v.V[(long long)i * n + row] : 0.0;
        double s = v.x[row];
#pragma unroll
        for (int i = 0; i < kMaxKrylov; ++i)
            if (i < k) s += ys[i] * vv[i];
        v.x[row] = s;
    }
}

}  // namespace sfh
