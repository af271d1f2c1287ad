// gmres.hip -- FP64 GMRES for M mu = b over the stored dipole matrix, all workgroups of the GPU on
// one or a few large surfaces (the north-star alternative to the LU of MatrixSolver::solve,
// CuSuperHelium/MatrixSolver.cuh:113-172).
//
// M = I/2 + K is a second-kind operator: its spectrum clusters at 1/2 (for the helium film one
// outlier sits far to the right), so GMRES converges in a handful of iterations, and the stepper
// warm starts every solve from the previous stage's mu.
//
// A solve is a fixed sequence of L launches of one kernel; launch i first finishes Arnoldi step
// i - 1 and then, unless the solve is done, runs step i.  No workgroup waits for another: every
// grid-wide result is read by the next launch.
//   finish step i-1 (i >= 1), in every workgroup: sum the partial dots v_l . w_{i-1} of all
//     workgroups (fixed order, so every workgroup gets the same h_{l,i-1}); rebuild
//     u = w_{i-1} - sum_l h_{l,i-1} v_l over all N entries into LDS (classical Gram-Schmidt, one
//     pass) and its norm h_{i,i-1} = |u|; rotate the Hessenberg column by the stored Givens
//     rotations and form the new one.  Workgroup 0 stores the column, the rotation and the scale.
//     Each workgroup stores its own rows of V[i] = u.
//   converged (or the last launch): y from the rotated Hessenberg system; each workgroup applies
//     x += V y to its own rows and pushes them into the warm-start history.
//   step i: w_i = M v_i for the workgroup's rows (one row per wave, the row's loads issued at the
//     kernel's start so they overlap the finishing work), and the partial dots v_l . w_i, l <= i.
// Launch 0 takes v_0 = r0 / |r0|.  The assembly kernel forms r0 = b - M x0 while it writes M
// (assemble.hip), and every workgroup of launch 0 reduces |r0| and |b| over its own full copies.
// L launches therefore run up to L - 1 Arnoldi steps.  Once the residual estimate reaches
// tol * |b|, the remaining launches return at once.  Partial dots and the rotated right-hand side
// live in per-launch (even/odd or per-index) slots, because a slot must not be overwritten while
// a slower workgroup of the same launch still reads it.
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
                                        double* red4, const double* pre = nullptr,
                                        const double* v0pre = nullptr) {
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
                    vv[q][e] = (v0pre && base == 0 && i + q == 0) ? v0pre[e]
                                                                   : (c < n ? V[(long long)(i + q) * n + c] : 0.0);
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
                vv[e] = (v0pre && base == 0 && i == 0) ? v0pre[e] : (c < n ? V[(long long)i * n + c] : 0.0);
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

// Rotates column jc of the Hessenberg matrix (entries 0..jc+1 in col, LDS) by the stored rotations
// 0..jc-1 and forms rotation jc; returns it in (c, s).
__device__ inline void givens_column(double* col, const double* cs, const double* sn, int jc, double& c,
                                     double& s) {
    for (int q = 0; q < jc; ++q) {
        const double t0 = cs[q] * col[q] + sn[q] * col[q + 1];
        const double t1 = -sn[q] * col[q] + cs[q] * col[q + 1];
        col[q] = t0;
        col[q + 1] = t1;
    }
    const double den = hypot(col[jc], col[jc + 1]);
    c = den > 0.0 ? col[jc] / den : 1.0;
    s = den > 0.0 ? col[jc + 1] / den : 0.0;
    col[jc] = den;
    col[jc + 1] = 0.0;
}

// This surface's history of solutions of the current RK stage (null for a plain evaluation), and
// the stage's count of stored steps (one increment per solve, by surface 0).
__device__ inline double* history(const GmArgs& a, const SurfDesc* sd) {
    return a.stage > 0 ? a.hist + (long long)(a.stage - 1) * kHistDepth * a.total + sd[blockIdx.z].voff : nullptr;
}
__device__ inline void count_history(const GmArgs& a) {
    a.hvalid[a.stage - 1] = min(a.hvalid[a.stage - 1] + 1, kHistDepth);
}

// Launch i of a solve's sequence (see the file comment).
__global__ __launch_bounds__(256) void gm_iter_kernel(GmArgs a, const SurfDesc* sd, int i) {
    GmCtl* ctl = reinterpret_cast<GmCtl*>(a.ctl) + blockIdx.z;
    if (ctl->done) return;   // converged solves: one load and out
    const GmView v = gm_view(a, sd);
    const int m = a.m;
    // the last launch of the sequence runs no product: it takes kCloseRows rows per workgroup
    // (fewer workgroups repeating the finishing work); the others kRowsPerWG
    const int rpw = (i == m - 1) ? kCloseRows : kRowsPerWG;
    if ((int)blockIdx.x * rpw >= v.n) return;
    __shared__ double vs[kGmMaxN];
    __shared__ double co[kMaxKrylov + 1];
    __shared__ double col[kMaxKrylov + 2];
    __shared__ double red8[8];
    __shared__ double wl[kRowsPerWG];
    __shared__ double sums[kMaxKrylov + 2];
    __shared__ double ys[kMaxKrylov + 4];   // g[j], gc[j+1], c, s, then y
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, n = v.n;
    const int r0 = blockIdx.x * rpw;
    const int row = r0 + wv;
    const double* R = v.M + (long long)min(row, n - 1) * n;
    const long long pslot = (long long)v.nwg * (m + 1);   // one partial-dot slot per launch parity
    RowChunk rc;
    if (i < m - 1) row_load(R, n, lane, rc);   // independent of v_i: in flight first
    int k = 0;        // Arnoldi steps completed
    double sj = 0.0;  // scale of v_i = sj * vs
    double sc_t = 0.0;
    bool done = false;
    // this workgroup's own rows (threads t < rpw): x, V[0] and the history levels that move
    // down, for the correction at the end of a solve, loaded up front
    double* h0 = history(a, sd);
    const int oc = min(r0 + t, n - 1);
    double xo = 0.0, vo0 = 0.0, ho[kHistDepth - 1];
#pragma unroll
    for (int l = 0; l < kHistDepth - 1; ++l) ho[l] = 0.0;
    if (t < rpw) {
        xo = v.x[oc];
        vo0 = v.V[oc];
        if (h0) {
#pragma unroll
            for (int l = 0; l < kHistDepth - 1; ++l) ho[l] = h0[(long long)l * a.total + oc];
        }
    }
    if (i == 0) {
        // v_0 = r0 / |r0|; beta = |r0| and |b| reduced by every workgroup over its own copies
        double sq = 0.0, bq = 0.0;
        for (int c = t; c < n; c += 256) {
            const double x = v.V[c];
            vs[c] = x;
            sq += x * x;
            const double bc = v.b[c];
            bq += bc * bc;
        }
        sq = wave_sum(sq);
        bq = wave_sum(bq);
        if (lane == 0) {
            red8[wv] = sq;
            red8[4 + wv] = bq;
        }
        __syncthreads();
        const double beta = sqrt(red8[0] + red8[1] + red8[2] + red8[3]);
        const double bnorm = sqrt(red8[4] + red8[5] + red8[6] + red8[7]);
        if (blockIdx.x == 0 && t == 0) {
            ctl->bnorm = bnorm;
            ctl->beta = beta;
            ctl->res = beta;
            v.gc[0] = beta;
            v.sc[0] = beta > 0.0 ? 1.0 / beta : 0.0;
        }
        done = beta <= a.tol * bnorm || m < 2;
        if (done && blockIdx.x == 0 && t == 0 && !(beta <= a.tol * bnorm)) ctl->unconverged += 1;
        sj = beta > 0.0 ? 1.0 / beta : 0.0;
    } else {
        // ---- finish step i - 1
        const int j = i - 1;
        const double* part = v.part + (j & 1) * pslot;
        double wpre[8];
        const double* wprev = (j & 1) ? v.w1 : v.w0;
        // loads that do not depend on the dot sums go out first, together with the partial dots
        double v0pre[8];
        load_first_pass(wprev, n, wpre);
        load_first_pass(v.V, n, v0pre);
        sc_t = t <= j ? v.sc[t] : 0.0;
        const double gold = v.gc[j];
        const double bnorm = ctl->bnorm;
        reduce_partials(part, v.nwg, j + 1, sums);
        if (t <= j) co[t] = sums[t] * sc_t;
        __syncthreads();
        const double hn = sqrt(rebuild_vector(wprev, v.V, co, j + 1, n, 0, vs, nullptr, red8, wpre, v0pre));
        if (t == 0) {
            for (int q = 0; q <= j; ++q) col[q] = sums[q];
            col[j + 1] = hn;
            double c, s;
            givens_column(col, v.cs, v.sn, j, c, s);
            ys[0] = c * gold;          // g[j]
            ys[1] = -s * gold;         // gc[j + 1]
            ys[2] = c;
            ys[3] = s;
        }
        __syncthreads();
        const double gj = ys[0], gnext = ys[1], res = fabs(gnext);
        if (blockIdx.x == 0 && t == 0) {
            double* Hc = v.H + (long long)j * (m + 1);
            for (int q = 0; q <= j + 1; ++q) Hc[q] = col[q];
            v.cs[j] = ys[2];
            v.sn[j] = ys[3];
            v.g[j] = gj;
            v.gc[j + 1] = gnext;
            v.sc[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
            ctl->res = res;
            ctl->hnext = hn;
        }
        k = i;
        const bool conv = res <= a.tol * bnorm || hn == 0.0;
        done = conv || i >= m - 1;
        if (done && !conv && blockIdx.x == 0 && t == 0) ctl->unconverged += 1;
        if (!done) {
            // V[i] = u: this workgroup's rows (every workgroup holds all of u in LDS)
            if (t < kRowsPerWG && r0 + t < n) v.V[(long long)i * n + r0 + t] = vs[r0 + t];
            sj = hn > 0.0 ? 1.0 / hn : 0.0;
        } else if (t == 0) {
            // y from R y = g over the k rotated columns (column k-1 is still in LDS), scaled by sc
            for (int q = k - 1; q >= 0; --q) {
                double acc = (q == k - 1) ? gj : v.g[q];
                for (int l = q + 1; l < k; ++l) {
                    const double hql = (l == k - 1) ? col[q] : v.H[(long long)l * (m + 1) + q];
                    acc -= hql * ys[4 + l];
                }
                const double hqq = (q == k - 1) ? col[q] : v.H[(long long)q * (m + 1) + q];
                ys[4 + q] = acc / hqq;
            }
        }
        if (done) {   // y_q sc_q
            __syncthreads();
            if (t <= j) ys[4 + t] *= sc_t;
        }
    }
    if (done) {
        // x += V y on this workgroup's rows; the stage's history
        __syncthreads();
        if (t < rpw && r0 + t < n) {
            const int c = r0 + t;
            double x = xo;
            if (k > 0) x += ys[4] * vo0;
            for (int q = 1; q < k; ++q) x += ys[4 + q] * v.V[(long long)q * n + c];
            v.x[c] = x;
            if (h0) {
#pragma unroll
                for (int l = kHistDepth - 1; l > 0; --l) h0[(long long)l * a.total + c] = ho[l - 1];
                h0[c] = x;
            }
        }
        if (blockIdx.x == 0 && t == 0) {
            ctl->iters = k;
            ctl->done = 1;
            if (h0 && blockIdx.z == 0) count_history(a);
        }
        return;
    }
    // ---- step i: w_i = M v_i (v_i = sj * vs) for this workgroup's rows, partial dots v_l . w_i
    double vrow[kRowsPerWG];
#pragma unroll
    for (int r = 0; r < kRowsPerWG; ++r)
        vrow[r] = (t < i && r0 + r < n) ? v.V[(long long)t * n + r0 + r] : 0.0;
    const double sct = (t < i) ? v.sc[t] : 0.0;
    double* wi = (i & 1) ? v.w1 : v.w0;
    const double d = row < n ? sj * row_dot_rest(R, rc, vs, n, lane) : 0.0;
    if (lane == 0) {
        wl[wv] = row < n ? d : 0.0;
        if (row < n) wi[row] = d;
    }
    __syncthreads();
    if (t <= i) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < kRowsPerWG; ++r) {
            const int rw = r0 + r;
            if (rw < n) {
                const double vi = (t == i) ? sj * vs[rw] : sct * vrow[r];
                s += vi * wl[r];
            }
        }
        v.part[(i & 1) * pslot + (long long)t * v.nwg + blockIdx.x] = s;
    }
}

}  // namespace sfh
