// gmres.hip -- FP64 GMRES for M mu = b over the stored dipole matrix, all workgroups of the GPU on
// one or a few large surfaces (the iterative alternative to the LU of MatrixSolver::solve,
// CuSuperHelium/MatrixSolver.cuh:113-172).
//
// M = I/2 + K is a second-kind operator: its spectrum clusters at 1/2 (for the helium film one
// outlier sits far to the right), so GMRES converges in a handful of iterations, and the stepper
// warm starts every solve from the previous stage's mu.
//
// A solve is a fixed sequence of L launches; launch i first finishes Arnoldi step i - 1 and then,
// unless the solve is done, runs step i.  No workgroup waits for another: every grid-wide result
// is read by the next launch.  Launches 0 .. L-2 are gm_iter_kernel; the last one, which only
// finishes a step, is gm_close_mu_kernel, which goes on to form the derivative mu' of the
// solution (with L = 1 the single launch is gm_iter_kernel).
//   finish step i-1 (i >= 1), in every workgroup: sum the partial dots v_l . w_{i-1} of all
//     workgroups (fixed order, so every workgroup gets the same h_{l,i-1}); rebuild
//     u = w_{i-1} - sum_l h_{l,i-1} v_l over all N entries into LDS (classical Gram-Schmidt, one
//     pass) and its norm h_{i,i-1} = |u|; rotate the Hessenberg column by the stored Givens
//     rotations and form the new one.  Workgroup 0 stores the column, the rotation and the scale.
//     Each workgroup stores its own rows of V[i] = u.
//   converged (or the last launch): y from the rotated Hessenberg system; the last launch stores
//     x0 + V y as the solution, each workgroup its own rows, and pushes them into the warm-start
//     history (a launch that finds convergence earlier only keeps y for it).  The initial guess
//     x0 and the solution are separate arrays.
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
// polynomial through them (extrapolate_history; with five steps kept its error is fifth order in
// dt, against first order for the previous stage's mu), which saves Krylov steps on every solve of
// a time integration.
#include "sfh_gmres.hpp"

namespace sfh {

// Row R (global) times vs (LDS) by one wave, in two halves so the first 2048 columns' loads can
// be issued before vs is ready: row_load issues 32 loads per lane (predicated at the row end),
// row_dot_rest finishes the product (columns beyond 2048 are loaded and used chunk by chunk).
// The matrix is FP64, or FP32 in the mixed-precision form (entries widened to FP64 as they are used).
template <typename T>
struct RowChunk {
    T mv[32];
};
template <typename T>
__device__ inline void row_load(const T* __restrict__ R, int n, int lane, RowChunk<T>& rc) {
#pragma unroll
    for (int u = 0; u < 32; ++u) {
        const int c = lane + 64 * u;
        rc.mv[u] = c < n ? R[c] : T(0);
    }
}
template <typename T>
__device__ inline double row_dot_rest(const T* __restrict__ R, const RowChunk<T>& rc, const double* vs, int n,
                                      int lane) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int u = 0; u < 32; u += 2) {
        const int c = lane + 64 * u;
        a0 += (double)rc.mv[u] * (c < n ? vs[c] : 0.0);
        a1 += (double)rc.mv[u + 1] * (c + 64 < n ? vs[c + 64] : 0.0);
    }
    for (int c0 = 32 * 64; c0 < n; c0 += 32 * 64) {
        T mv[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int c = c0 + lane + 64 * u;
            mv[u] = c < n ? R[c] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 32; u += 2) {
            const int c = c0 + lane + 64 * u;
            a0 += (double)mv[u] * (c < n ? vs[c] : 0.0);
            a1 += (double)mv[u + 1] * (c + 64 < n ? vs[c + 64] : 0.0);
        }
    }
    return wave_sum(a0 + a1);
}

// u = w - sum_{i<j} co[i] V[i] over all n entries into vs (LDS, unless null); returns |u|^2 over
// the workgroup.
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
                if (vs) vs[c] = u[e];
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

// y from R y = g over the k rotated columns, into y[0..k): column k - 1 (col, LDS) is this
// launch's, columns before it (H, column l at H[l (m+1)]) and g[0..k-2] come from earlier
// launches.
__device__ inline void back_substitute(const double* g, const double* H, int m, int k, double gk, const double* col,
                                       double* y) {
    for (int q = k - 1; q >= 0; --q) {
        double acc = (q == k - 1) ? gk : g[q];
        for (int l = q + 1; l < k; ++l) {
            const double hql = (l == k - 1) ? col[q] : H[l * (m + 1) + q];
            acc -= hql * y[l];
        }
        const double hqq = (q == k - 1) ? col[q] : H[q * (m + 1) + q];
        y[q] = acc / hqq;
    }
}

// Launch i of a solve's sequence (see the file comment).
template <typename T>
__global__ __launch_bounds__(256) void gm_iter_kernel(GmArgs a, const SurfDesc* __restrict__ sd, int i) {
    GmCtl* ctl = reinterpret_cast<GmCtl*>(a.ctl) + blockIdx.z;
    // solves finished by an earlier launch return at once.  A finish stamped by this launch
    // (workgroup 0 can get there before a late workgroup starts) does not count.  Launch 0 needs
    // no check (the derivative kernel re-armed the block), and launch 1 makes it after its row
    // loads have gone out (an earlier finish is the exception there); later launches first.
    if (i >= 2 || (i == 1 && a.check_first)) {
        const int fin = ctl->done;
        if (fin != 0 && fin <= i) return;
    }
    const GmView v = gm_view(a, sd);
    const int m = a.m;
    const int rpw = kRowsPerWG;
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
    const T* R = (sizeof(T) == 8 ? reinterpret_cast<const T*>(v.M) : reinterpret_cast<const T*>(v.M32)) +
                 (long long)min(row, n - 1) * n;
    const long long pslot = (long long)v.nwg * (m + 1);   // one partial-dot slot per launch parity
    // Loads go out in the order their results are needed, because a wait for one load waits for
    // all loads issued before it: first the vectors and sums of the step being finished (launch
    // 0: r0 and b), then the matrix rows, which the finishing work then overlaps.
    const int j = i - 1;
    const double* wprev = (j & 1) ? v.w1 : v.w0;
    const double* part = v.part + (j & 1) * pslot;
    double pa[8], pb[8];   // first 2048 entries: launch 0 r0 and b, later w_{i-1} and V[0]
    DotLoads dl;
    const bool fastdots = i >= 1 && dots_fast(v.nwg, i);
    double sc_t = 0.0, gold = 0.0, bnorm = 0.0;
    if (i == 0) {
        load_first_pass(v.V, n, pa);
        load_first_pass(v.b, n, pb);
    } else {
        load_first_pass(wprev, n, pa);
        load_first_pass(v.V, n, pb);
        if (fastdots) dots_load(part, v.nwg, i, dl);
        sc_t = t <= j ? v.sc[t] : 0.0;
        gold = v.gc[j];
        bnorm = ctl->bnorm;
    }
    // the rows up front, except in a refinement solve, which most often ends at once: there they
    // are fetched when the step is known to run
    RowChunk<T> rc;
    const bool early_rows = !a.check_first;
    if (i < m - 1 && early_rows) row_load(R, n, lane, rc);
    if (i == 1 && !a.check_first) {
        const int fin = ctl->done;
        if (fin != 0 && fin <= i) return;
    }
    int k = 0;        // Arnoldi steps completed
    double sj = 0.0;  // scale of v_i = sj * vs
    bool done = false;
    // this workgroup's own rows (threads t < rpw): x, V[0] and the history levels that move
    // down, for the correction at the end of a solve, loaded up front
    double* h0 = history(a, sd, blockIdx.z);
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
        for (int base = 0; base < n; base += 2048) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = base + t + 256 * e;
                const double x = base == 0 ? pa[e] : (c < n ? v.V[c] : 0.0);
                const double bc = base == 0 ? pb[e] : (c < n ? v.b[c] : 0.0);
                if (c < n) {
                    vs[c] = x;
                    sq += x * x;
                    bq += bc * bc;
                }
            }
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
        if (fastdots)
            dots_finish(dl, j + 1, sums);
        else
            reduce_partials(part, v.nwg, j + 1, sums);
        if (t <= j) co[t] = sums[t] * sc_t;
        __syncthreads();
        const double hn = sqrt(rebuild_vector(wprev, v.V, co, j + 1, n, 0, vs, nullptr, red8, pa, pb));
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
            back_substitute(v.g, v.H, m, k, gj, col, ys + 4);   // then scaled by sc
        }
        if (done) {   // y_q sc_q
            __syncthreads();
            if (t <= j) ys[4 + t] *= sc_t;
        }
    }
    if (done && i < m - 1) {
        // converged before the last launch, which stores the solution (gm_close_mu_kernel): keep y
        if (blockIdx.x == 0) {
            if (t < k) v.y[t] = ys[4 + t];   // scaled by this thread above
            if (t == 0) {
                set_iters(ctl, a, k);
                ctl->done = i + 1;
            }
        }
        return;
    }
    if (done) {
        // the single launch of a one-launch sequence: x0 + V y on this workgroup's rows; the
        // stage's history
        __syncthreads();
        if (t < rpw && r0 + t < n) {
            const int c = r0 + t;
            double x = xo;
            if (k > 0) x += ys[4] * vo0;
            for (int q = 1; q < k; ++q) x += ys[4 + q] * v.V[(long long)q * n + c];
            v.sol[c] = x;
            if (h0) {
#pragma unroll
                for (int l = kHistDepth - 1; l > 0; --l) h0[(long long)l * a.total + c] = ho[l - 1];
                h0[c] = x;
            }
        }
        if (blockIdx.x == 0 && t == 0) {
            set_iters(ctl, a, k);
            ctl->done = i + 1;
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
    if (!early_rows) row_load(R, n, lane, rc);
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
    if (i == 0 && t == 1) {
        // launch 0 also leaves |w_0|^2 (slot 1): a sequence of one Arnoldi step is finished from
        // the two sums alone by velocity_gm1_kernel
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < kRowsPerWG; ++r) s += wl[r] * wl[r];
        v.part[(long long)v.nwg + blockIdx.x] = s;
    }
}

// The last launch of a sequence (i = m - 1 >= 1) fused with the derivative of the solution,
// mu' = D1 mu 2 pi / n (mu_circ_kernel's circulant product; BaseBoundaryIntegrator.cuh:201-203).
// Every workgroup finishes Arnoldi step i - 1 as gm_iter_kernel does, forms y and stages
// mu = x0 + V y of all nodes in LDS, kCircChunk at a time.  It stores its own kCircRowsPerWG rows
// of the solution (and pushes them into the stage's history) from the staged values, and
// multiplies the staged mu by the circulant kernel for mu' of the same rows.  A solve that
// converged in an earlier launch left its (scaled) y behind, and this launch stores its solution
// all the same.  Everything the launch reads that does not depend on another of its loads is
// requested at the start: the kernel window of the first chunk, the first 2048 entries of x0,
// V[0], V[1] and w_{i-1}, the partial dots, and the small state (scales, rotations, g, H, y) into
// LDS, so that the finishing work runs on data already at hand.  A workgroup reads the initial
// guess, the basis and earlier launches' results, never the solution entries that other
// workgroups of this launch are storing.  grid = (ceil(max n / kCircRowsPerWG), surfaces).
__global__ __launch_bounds__(256) void gm_close_mu_kernel(GmArgs a, const SurfDesc* __restrict__ sd,
                                                          const double* __restrict__ circ,
                                                          double2* __restrict__ mup, int i) {
    constexpr int CR = kCircRowsPerWG;
    constexpr int KM = kMaxKrylov;
    static_assert(kCircPer == 8, "the first staged chunk reuses the 8-entry first passes");
    const int s = blockIdx.y;
    const SurfDesc d = sd[s];
    const int n = d.n;
    const int k0 = blockIdx.x * CR;
    if (k0 >= n) return;
    const GmView v = gm_view(a, sd, s);
    GmCtl* ctl = reinterpret_cast<GmCtl*>(a.ctl) + s;
    const int m = a.m;
    const int j = i - 1;
    __shared__ double ms[kCircChunk];
    __shared__ double2 c1[kCircWin];
    __shared__ double Hs[KM * (KM + 1)];
    __shared__ double cs_s[KM], sn_s[KM], g_s[KM + 1];
    __shared__ double co[KM + 1];
    __shared__ double col[KM + 2];
    __shared__ double red8[8];
    __shared__ double sums[KM + 2];
    __shared__ double ys[KM + 4];   // g[j], gc[j+1], c, s, then y
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const double* c1r = circ + d.coff;
    const double* c1i = c1r + n;
    double ar[kCircWinPer], ai[kCircWinPer];
    {
        int q[kCircWinPer];
        circ_window(k0, 0, n, q);
#pragma unroll
        for (int e = 0; e < kCircWinPer; ++e) {
            ar[e] = c1r[q[e]];
            ai[e] = c1i[q[e]];
        }
    }
    const double* wprev = (j & 1) ? v.w1 : v.w0;
    double xpre[8], v0pre[8], v1pre[8], wpre[8];
    load_first_pass(v.x, n, xpre);
    load_first_pass(v.V, n, v0pre);
    load_first_pass(j >= 1 ? v.V + n : v.V, n, v1pre);
    load_first_pass(wprev, n, wpre);
    const long long pslot = (long long)v.nwg * (m + 1);
    const double* part = v.part + (j & 1) * pslot;
    DotLoads dl;
    const bool fastdots = dots_fast(v.nwg, j + 1);
    if (fastdots) dots_load(part, v.nwg, j + 1, dl);
    const double sc_t = t <= j ? v.sc[t] : 0.0;
    const double y_t = t < m ? v.y[t] : 0.0;
    const double gold = v.gc[j];
    double cs_t = 0.0, sn_t = 0.0, g_t = 0.0, hreg[4];
    if (t < j) {
        cs_t = v.cs[t];
        sn_t = v.sn[t];
        g_t = v.g[t];
    }
    static_assert(4 * 256 >= (KM - 1) * (KM + 1), "H columns 0 .. j-1 in four registers per thread");
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = t + 256 * r;
        hreg[r] = e < j * (m + 1) ? v.H[e] : 0.0;
    }
    double* h0 = history(a, sd, s);
    double ho[kHistDepth - 1];
#pragma unroll
    for (int l = 0; l < kHistDepth - 1; ++l) ho[l] = 0.0;
    if (h0 && t < CR) {
        const int oc = min(k0 + t, n - 1);
#pragma unroll
        for (int l = 0; l < kHistDepth - 1; ++l) ho[l] = h0[(long long)l * a.total + oc];
    }
    // the control block last: its fields decide branches, so the first use waits for them
    const int fin = ctl->done;
    const int iters_before = a.check_first ? ctl->refine_iters : ctl->iters;
    const double bnorm = ctl->bnorm;
    // all loads are out; the small state into LDS and the partial dots summed (used when the
    // solve is open)
    if (t < j) {
        cs_s[t] = cs_t;
        sn_s[t] = sn_t;
        g_s[t] = g_t;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = t + 256 * r;
        if (e < j * (m + 1)) Hs[e] = hreg[r];
    }
    if (fastdots)
        dots_finish(dl, j + 1, sums);
    else
        reduce_partials(part, v.nwg, j + 1, sums);
    // a finish stamped by this launch (workgroup 0 got there first) leaves the solve open here
    const bool open = !(fin != 0 && fin <= i);
    int k;
    if (!open) {
        // converged in an earlier launch, which kept y (scaled) for this one
        k = iters_before;
        if (t < k) ys[4 + t] = y_t;
        __syncthreads();
    } else {
        // ---- finish step j: h_{l,j} = sums[l] sc_l, |u| with u = w_j - sum_l h_{l,j} v_l
        if (t <= j) co[t] = sums[t] * sc_t;
        __syncthreads();
        double ss = 0.0;
        for (int base = 0; base < n; base += 2048) {
            double u[8];
            if (base == 0) {
#pragma unroll
                for (int e = 0; e < 8; ++e) u[e] = wpre[e] - co[0] * v0pre[e];
                if (j >= 1) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) u[e] -= co[1] * v1pre[e];
                }
                for (int l = 2; l <= j; ++l) {
                    double vv[8];
                    load_first_pass(v.V + (long long)l * n, n, vv);
#pragma unroll
                    for (int e = 0; e < 8; ++e) u[e] -= co[l] * vv[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int c = base + t + 256 * e;
                    u[e] = c < n ? wprev[c] : 0.0;
                }
                for (int l = 0; l <= j; ++l) {
                    double vv[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int c = base + t + 256 * e;
                        vv[e] = c < n ? v.V[(long long)l * n + c] : 0.0;
                    }
#pragma unroll
                    for (int e = 0; e < 8; ++e) u[e] -= co[l] * vv[e];
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = base + t + 256 * e;
                if (c < n) ss += u[e] * u[e];
            }
        }
        const double hn = sqrt(block_sum4(ss, red8));
        if (t == 0) {
            for (int q = 0; q <= j; ++q) col[q] = sums[q];
            col[j + 1] = hn;
            double c, sn;
            givens_column(col, cs_s, sn_s, j, c, sn);
            ys[0] = c * gold;    // g[j]
            ys[1] = -sn * gold;  // gc[j + 1]
            ys[2] = c;
            ys[3] = sn;
            back_substitute(g_s, Hs, m, j + 1, ys[0], col, ys + 4);
        }
        __syncthreads();
        const double gj = ys[0], gnext = ys[1], res = fabs(gnext);
        k = i;
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
            if (!(res <= a.tol * bnorm || hn == 0.0)) ctl->unconverged += 1;
            set_iters(ctl, a, k);
            ctl->done = i + 1;
        }
        if (t <= j) ys[4 + t] *= sc_t;   // y_q sc_q
        __syncthreads();
    }
    if (h0 && s == 0 && blockIdx.x == 0 && t == 0) count_history(a);
    // mu = x0 + y_0 V[0] + y_1 V[1] (+ the rest when k > 2), without branches in the common case
    const double y0 = k > 0 ? ys[4] : 0.0, y1 = k > 1 ? ys[5] : 0.0;
    double2 a0s = {0.0, 0.0}, a1s = {0.0, 0.0};
    for (int j0 = 0; j0 < n; j0 += kCircChunk) {
        const int jn = min(kCircChunk, n - j0);
        if (j0 > 0) {
            __syncthreads();   // the previous chunk has been consumed
            int q[kCircWinPer];
            circ_window(k0, j0, n, q);
#pragma unroll
            for (int e = 0; e < kCircWinPer; ++e) {
                ar[e] = c1r[q[e]];
                ai[e] = c1i[q[e]];
            }
            load_first_pass(v.x + j0, n - j0, xpre);
            load_first_pass(v.V + j0, n - j0, v0pre);
            load_first_pass(v.V + n + j0, n - j0, v1pre);
        }
        double mv[kCircPer];
#pragma unroll
        for (int e = 0; e < kCircPer; ++e) mv[e] = xpre[e] + y0 * v0pre[e] + y1 * v1pre[e];
        for (int q = 2; q < k; ++q) {
            double vq[8];
            load_first_pass(v.V + (long long)q * n + j0, n - j0, vq);
            const double yq = ys[4 + q];
#pragma unroll
            for (int e = 0; e < kCircPer; ++e) mv[e] += yq * vq[e];
        }
#pragma unroll
        for (int e = 0; e < kCircPer; ++e) {
            const int c = t + 256 * e;
            if (c < jn) ms[c] = mv[e];
        }
#pragma unroll
        for (int e = 0; e < kCircWinPer; ++e) {
            const int u = t + 256 * e;
            if (u < kCircWin) c1[u] = make_double2(ar[e], ai[e]);
        }
        __syncthreads();
        if (t < CR) {   // the own rows of the solution, as staged
            const int r = k0 + t;
            if (r < n && r >= j0 && r < j0 + jn) {
                const double x = ms[r - j0];
                v.sol[r] = x;
                if (h0) {
#pragma unroll
                    for (int l = kHistDepth - 1; l > 0; --l) h0[(long long)l * a.total + r] = ho[l - 1];
                    h0[r] = x;
                }
            }
        }
#pragma unroll 4
        for (int c = lane; c < jn; c += 64) {
            const double x = ms[c];
            const int u0 = wv - c + kCircChunk - 1;
            const double2 p0 = c1[u0], p1 = c1[u0 + 4];
            a0s.x += p0.x * x;
            a0s.y += p0.y * x;
            a1s.x += p1.x * x;
            a1s.y += p1.y * x;
        }
    }
    a0s.x = wave_sum(a0s.x); a0s.y = wave_sum(a0s.y);
    a1s.x = wave_sum(a1s.x); a1s.y = wave_sum(a1s.y);
    if (lane < 2) {   // lane 0: row wv, lane 1: row wv + 4
        const int r = k0 + wv + 4 * lane;
        if (r < n) {
            const double2 acc = lane ? a1s : a0s;
            const double xs = 2.0 * kPi / n;
            mup[d.voff + r] = make_double2(acc.x * xs, acc.y * xs);
        }
    }
}

template __global__ void gm_iter_kernel<double>(GmArgs, const SurfDesc*, int);
template __global__ void gm_iter_kernel<float>(GmArgs, const SurfDesc*, int);

}  // namespace sfh
