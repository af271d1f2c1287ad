// gmres.hip -- FP64 GMRES for M mu = b over the stored dipole matrix, all workgroups of the GPU on
// one or a few large surfaces (the north-star alternative to the LU of MatrixSolver::solve,
// CuSuperHelium/MatrixSolver.cuh:113-172).
//
// M = I/2 + K is a second-kind operator: its spectrum clusters at 1/2 (for the helium film one
// outlier sits far to the right), so GMRES converges in a handful of iterations, and the stepper
// warm starts every solve from the previous stage's mu.
//
// ONE kernel per Arnoldi step.  Kernel j:
//   1. loads v_j = sc[j] V[j] into LDS and multiplies its rows of M by it (one row per wave, the
//      whole row's loads in flight), writing w_j;
//   2. forms partial dots v_i . w_j (i <= j) of its rows;
//   3. the last workgroup (release increments, one acquire) reduces the dots, forms the next
//      basis vector u = w_j - sum_i h_{i,j} v_i over all N entries (classical Gram-Schmidt, one
//      pass), its norm h_{j+1,j} = |u|, stores V[j+1] = u and sc[j+1] = 1 / |u|, applies the
//      Givens rotations to column j and tests the residual estimate.
// The initial residual r0 = b - M x0 is formed by the assembly kernel while it writes M
// (assemble.hip) and normalised by step 0 (every workgroup reduces |r0| and |b| over its own
// full copies, so no grid-wide reduction precedes the first product), and the last workgroup of
// the final step applies the correction x += V y, so a solve costs the Arnoldi steps only.  Once the residual estimate reaches
// tol * |b| the remaining launches of the fixed-length sequence return at once.
//
// Initial guess of an RK stage: mu of the same stage extrapolated over the previous steps by the
// polynomial through them (extrapolate_history; with four steps kept its error is fourth order in
// dt, against first order for the previous stage's mu), which saves Krylov steps on every solve of
// a time integration.
#include "sfh_gmres.hpp"

namespace sfh {

// Row R (global) times vs (LDS) by one wave, in two halves so the first 2048 columns' loads can
// be issued before vs is ready: row_load issues 32 loads per lane (predicated at the row end),
// row_dot_rest finishes the product (columns beyond 2048 are loaded and used chunk by chunk).
struct RowChunk {
    double mv[32];
};
__device__ inline void row_load(const double* __restrict__ R, int n, int lane, RowChunk& rc) {
#pragma unroll
    for (int u = 0; u < 32; ++u) {
        const int c = lane + 64 * u;
        rc.mv[u] = c < n ? R[c] : 0.0;
    }
}
__device__ inline double row_dot_rest(const double* __restrict__ R, const RowChunk& rc, const double* vs, int n,
                                      int lane) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int u = 0; u < 32; u += 2) {
        const int c = lane + 64 * u;
        a0 += rc.mv[u] * (c < n ? vs[c] : 0.0);
        a1 += rc.mv[u + 1] * (c + 64 < n ? vs[c + 64] : 0.0);
    }
    for (int c0 = 32 * 64; c0 < n; c0 += 32 * 64) {
        double mv[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int c = c0 + lane + 64 * u;
            mv[u] = c < n ? R[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 32; u += 2) {
            const int c = c0 + lane + 64 * u;
            a0 += mv[u] * (c < n ? vs[c] : 0.0);
            a1 += mv[u + 1] * (c + 64 < n ? vs[c + 64] : 0.0);
        }
    }
    return wave_sum(a0 + a1);
}

// u = w - sum_{i<j} co[i] V[i] over all n entries into vs (LDS); returns |u|^2 over the workgroup.
// Each thread handles 8 entries per pass with the basis loads of 4 vectors in flight together.
// The entries of rows [r0, r0 + kRowsPerWG) are also stored into Vj (when not null).
__device__ inline void load_first_pass(const double* __restrict__ w, int n, double (&u)[8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = threadIdx.x + 256 * e;
        u[e] = c < n ? w[c] : 0.0;
    }
}
__device__ inline double rebuild_vector(const double* __restrict__ w, const double* __restrict__ V,
                                        const double* co, int j, int n, int r0, double* vs, double* Vj,
                                        double* red4, const double* pre = nullptr) {
    const int t = threadIdx.x;
    double ss = 0.0;
    for (int base = 0; base < n; base += 8 * 256) {
        double u[8];
        if (pre && base == 0) {
#pragma unroll
            for (int e = 0; e < 8; ++e) u[e] = pre[e];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = base + t + 256 * e;
                u[e] = c < n ? w[c] : 0.0;
            }
        }
        int i = 0;
        for (; i + 4 <= j; i += 4) {
            double vv[4][8];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int c = base + t + 256 * e;
                    vv[q][e] = c < n ? V[(long long)(i + q) * n + c] : 0.0;
                }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double cq = co[i + q];
#pragma unroll
                for (int e = 0; e < 8; ++e) u[e] -= cq * vv[q][e];
            }
        }
        for (; i < j; ++i) {
            double vv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = base + t + 256 * e;
                vv[e] = c < n ? V[(long long)i * n + c] : 0.0;
            }
            const double ci = co[i];
#pragma unroll
            for (int e = 0; e < 8; ++e) u[e] -= ci * vv[e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = base + t + 256 * e;
            if (c < n) {
                vs[c] = u[e];
                ss += u[e] * u[e];
                if (Vj && c >= r0 && c < r0 + kRowsPerWG) Vj[c] = u[e];
            }
        }
    }
    return block_sum4(ss, red4);
}

// This surface's history of solutions of the current RK stage (null for a plain evaluation), and
// the stage's count of stored steps (one increment per solve, by surface 0).
__device__ inline double* history(const GmArgs& a, const SurfDesc* sd) {
    return a.stage > 0 ? a.hist + (long long)(a.stage - 1) * kHistDepth * a.total + sd[blockIdx.z].voff : nullptr;
}
__device__ inline void count_history(const GmArgs& a) {
    a.hvalid[a.stage - 1] = min(a.hvalid[a.stage - 1] + 1, kHistDepth);
}

// Applies the stored rotations to column jc (entries 0..jc+1 in col), forms the new rotation and
// updates g; returns |g[jc+1]|, the residual estimate.
__device__ inline double givens_column(double* col, double* cs, double* sn, double* g, int jc) {
    for (int q = 0; q < jc; ++q) {
        const double t0 = cs[q] * col[q] + sn[q] * col[q + 1];
        const double t1 = -sn[q] * col[q] + cs[q] * col[q + 1];
        col[q] = t0;
        col[q + 1] = t1;
    }
    const double den = hypot(col[jc], col[jc + 1]);
    const double c = den > 0.0 ? col[jc] / den : 1.0;
    const double s = den > 0.0 ? col[jc + 1] / den : 0.0;
    cs[jc] = c;
    sn[jc] = s;
    col[jc] = den;
    col[jc + 1] = 0.0;
    g[jc + 1] = -s * g[jc];
    g[jc] = c * g[jc];
    return fabs(g[jc + 1]);
}

// Arnoldi step j (see the file comment).
__global__ __launch_bounds__(256) void gm_iter_kernel(GmArgs a, const SurfDesc* sd, int j) {
    // converged solves: one load and out (most launches of the fixed sequence end here)
    if (reinterpret_cast<const GmCtl*>(a.ctl)[blockIdx.z].done) return;
    const GmView v = gm_view(a, sd);
    if ((int)blockIdx.x >= v.nwg) return;
    __shared__ double vs[kGmMaxN];
    __shared__ double co[kMaxKrylov + 1];
    __shared__ double red4[4];
    __shared__ double wl[kRowsPerWG];
    __shared__ double sums[kMaxKrylov + 2];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, n = v.n, m = a.m;
    const int r0 = blockIdx.x * kRowsPerWG;
    double* wj = (j & 1) ? v.w1 : v.w0;
    const int row = r0 + wv;
    const double* R = v.M + (long long)min(row, n - 1) * n;
    RowChunk rc;
    row_load(R, n, lane, rc);   // the matrix row does not depend on v_j: in flight first
    // operands of the partial dots (v_i of this workgroup's rows, i < j), also independent of v_j
    double vrow[kRowsPerWG];
#pragma unroll
    for (int r = 0; r < kRowsPerWG; ++r)
        vrow[r] = (t < j && r0 + r < n) ? v.V[(long long)t * n + r0 + r] : 0.0;
    const double sct = (t < j) ? v.sc[t] : 0.0;
    // V[j] unscaled into LDS; step 0 also forms beta = |r0| and |b| (every workgroup, same order,
    // same values)
    double sq = 0.0, bq = 0.0;
    for (int c = t; c < n; c += 256) {
        const double x = v.V[(long long)j * n + c];
        vs[c] = x;
        sq += x * x;
        if (j == 0) {
            const double bc = v.b[c];
            bq += bc * bc;
        }
    }
    double beta = 0.0, bnorm = 0.0, sj;
    if (j == 0) {
        __shared__ double red8[8];
        sq = wave_sum(sq);
        bq = wave_sum(bq);
        if (lane == 0) {
            red8[wv] = sq;
            red8[4 + wv] = bq;
        }
        __syncthreads();
        beta = sqrt(red8[0] + red8[1] + red8[2] + red8[3]);
        bnorm = sqrt(red8[4] + red8[5] + red8[6] + red8[7]);
        if (beta <= a.tol * bnorm) {   // the initial guess already solves the system
            if (blockIdx.x == 0 && t == 0) {
                v.ctl->bnorm = bnorm;
                v.ctl->beta = beta;
                v.ctl->res = beta;
                v.ctl->done = 1;
            }
            double* h0 = history(a, sd);   // x = x0 enters the history, this workgroup's rows
            if (h0 && t < kRowsPerWG && r0 + t < n) push_history(h0 + r0 + t, a.total, v.x[r0 + t]);
            if (h0 && blockIdx.x == 0 && blockIdx.z == 0 && t == 0) count_history(a);
            return;
        }
        sj = 1.0 / beta;
    } else {
        sj = v.sc[j];
        __syncthreads();
    }
    // w_j = M v_j = sc_j (M V[j]) for this workgroup's rows
    const double d = row < n ? sj * row_dot_rest(R, rc, vs, n, lane) : 0.0;
    if (lane == 0) {
        wl[wv] = row < n ? d : 0.0;
        if (row < n) wj[row] = d;
    }
    __syncthreads();
    // partial dots v_i . w_j, i <= j
    if (t <= j) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < kRowsPerWG; ++r) {
            const int rw = r0 + r;
            if (rw < n) {
                const double vi = (t == j) ? sj * vs[rw] : sct * vrow[r];
                s += vi * wl[r];
            }
        }
        v.part[t * v.nwg + blockIdx.x] = s;
    }
    if (!last_workgroup(v.ctl, v.nwg)) return;
    double wpre[8];
    load_first_pass(wj, n, wpre);   // in flight together with the partial sums
    reduce_partials(v.part, v.nwg, j + 1, sums);
    if (t <= j) co[t] = sums[t] * (t == j ? sj : v.sc[t]);
    if (j == 0 && t == 0) {   // the solve's start: scale of v_0, the residual norm, |b|
        v.sc[0] = sj;
        v.g[0] = beta;
        v.ctl->beta = beta;
        v.ctl->bnorm = bnorm;
    }
    __syncthreads();
    // u = w_j - sum_{i<=j} h_{i,j} v_i, |u| (every entry of this one workgroup, into LDS)
    const double hn = sqrt(rebuild_vector(wj, v.V, co, j + 1, n, 0, vs, nullptr, red4, wpre));
    __shared__ int fin;
    if (t == 0) {
        GmCtl* c = v.ctl;
        double* col = v.H + (long long)j * (m + 1);
        for (int q = 0; q <= j; ++q) col[q] = sums[q];
        col[j + 1] = hn;
        v.sc[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
        c->res = givens_column(col, v.cs, v.sn, v.g, j);
        c->iters = j + 1;
        c->hnext = hn;
        if (c->res <= a.tol * c->bnorm || hn == 0.0) {
            c->done = 1;
        } else if (j + 1 >= m) {
            c->done = 1;
            c->unconverged += 1;
        }
        fin = c->done;
    }
    __syncthreads();
    if (!fin) {   // V[j+1] = u for the next step
        for (int c = t; c < n; c += 256) v.V[(long long)(j + 1) * n + c] = vs[c];
        return;
    }
    // the solve ends here: y from the rotated Hessenberg system R y = g (one thread), then
    // x += sum_i y_i v_i over all entries and the stage's history
    const int k = j + 1;
    if (t == 0) {
        for (int i = k - 1; i >= 0; --i) {   // co holds y first, then y_i sc[i]
            double acc = v.g[i];
            for (int l = i + 1; l < k; ++l) acc -= v.H[(long long)l * (m + 1) + i] * co[l];
            co[i] = acc / v.H[(long long)i * (m + 1) + i];
        }
        for (int i = 0; i < k; ++i) co[i] *= v.sc[i];
    }
    __syncthreads();
    double* h0 = history(a, sd);
    const long long T = a.total;
    for (int base = 0; base < n; base += 8 * 256) {   // 8 entries per thread, every load in flight
        double xs[8], ph[kHistDepth - 1][8];   // history levels 0 .. depth-2 move down one level
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = base + t + 256 * e;
            xs[e] = c < n ? v.x[c] : 0.0;
#pragma unroll
            for (int l = 0; l < kHistDepth - 1; ++l) ph[l][e] = (h0 && c < n) ? h0[l * T + c] : 0.0;
        }
        for (int i0 = 0; i0 < k; i0 += 4) {
            double vv[4][8];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int c = base + t + 256 * e;
                    vv[q][e] = (i0 + q < k && c < n) ? v.V[(long long)(i0 + q) * n + c] : 0.0;
                }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double y = i0 + q < k ? co[i0 + q] : 0.0;
#pragma unroll
                for (int e = 0; e < 8; ++e) xs[e] += y * vv[q][e];
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = base + t + 256 * e;
            if (c >= n) continue;
            v.x[c] = xs[e];
            if (h0) {
#pragma unroll
                for (int l = kHistDepth - 1; l > 0; --l) h0[l * T + c] = ph[l - 1][e];
                h0[c] = xs[e];
            }
        }
    }
    if (h0 && blockIdx.z == 0 && t == 0) count_history(a);
}

}  // namespace sfh
