// spectral.hip -- tangential derivatives of the multi-workgroup path as circulant products.
//
// Replaces ZPhiDerivative::exec (CuSuperHelium/Derivatives.cuh:310-384) and the mu' derivative of
// BaseBoundaryIntegrator.cuh:201-203 (FftDerivative::exec, Derivatives.cuh:189-257).  Forward
// DFT, multiplier table (utilities.cuh:119-147 first derivative, :185-199 second), inverse DFT is
// a circulant linear map of the samples:
//   y_k = sum_j c[(k - j) mod n] v_j,   c_q = sum_i m_i exp(+2 pi i i q / n),
// c being the inverse transform of the multipliers (complex for the first derivative, whose table
// is not conjugate-symmetric at n/2 and n/2 + 1; real for the second).  The tables are built once
// per node count in extended precision on the host.
//
// An in-LDS FFT of one field is a chain of dependent LDS passes inside a single workgroup, so at
// n = 2048 it takes ~4 us each way on one CU while the rest of the GPU idles.  As circulant
// products the three derivatives of a right-hand side (Z', Z'', Re Phi') are 7 n^2 multiply-adds
// spread over n / 8 workgroups of 8 rows; each workgroup stages the samples and the kernel
// window it needs in LDS, kCircChunk columns at a time with all of a chunk's loads in flight
// together, and every row is summed inside one workgroup (no cross-workgroup reduction).
#include "sfh_gmres.hpp"

namespace sfh {

constexpr int kCircRows = kCircRowsPerWG;   // rows per workgroup: wave w owns rows w and w + 4

struct CircLds {
    double2 z[kCircChunk];   // samples of the chunk (Z minus its linear part; mu in .x)
    double p[kCircChunk];    // phi minus its linear part
    double2 c1[kCircWin];    // first-derivative kernel over the chunk's window of k - j
    double c2[kCircWin];     // second-derivative kernel over the same window
};

// Z' (field scale 2 pi / n, plus the slope of the linear part), Z'' (4 pi^2 / n^2) and
// b = Re Phi' (2 pi / n, plus the uniform-flow term) for kCircRows rows of one surface;
// also the rows' half-angle table, and with GMRES the rows' initial guess and the solve's
// control block.  grid = (ceil(max n / kCircRows), surfaces).
__global__ __launch_bounds__(256) void zphi_circ_kernel(const SurfDesc* __restrict__ sd,
                                                        const double* __restrict__ circ,
                                                        const double2* __restrict__ Z,
                                                        const double* __restrict__ phi,
                                                        double2* __restrict__ Zp, double2* __restrict__ Zpp,
                                                        double* __restrict__ b, double* __restrict__ work,
                                                        Phys ph, int* __restrict__ status,
                                                        NodeTrig* __restrict__ trig, GmArgs gm, int gm_on) {
    __shared__ CircLds L;
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int k0 = blockIdx.x * kCircRows;
    if (k0 >= n) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // re-arm the surface's interchange flag / GMRES control block for the solve that follows
    if (blockIdx.x == 0 && t == 0) {
        if (status) status[2 * blockIdx.y] = 0;
        if (gm_on) {
            GmCtl* c = reinterpret_cast<GmCtl*>(gm.ctl) + blockIdx.y;
            c->done = 0;
            c->iters = 0;
            c->refine_iters = 0;
            if (gm.stage > 0 && gm.hvalid[gm.stage - 1] < kHistDepth) c->fill_iters = 0;
        }
    }
    const double* c1r = circ + d.coff;
    const double* c1i = c1r + n;
    const double* c2 = c1i + n;
    const double xs = 2.0 * kPi / n;                       // Derivatives.cuh:285,317
    const double plin = -(1 + ph.rho) * kPi * ph.U / n;    // Derivatives.cuh:286
    // the own rows' inputs of the half-angle table and of the initial guess (threads t < kCircRows),
    // loaded before the products so that their latency overlaps them
    const int kt = min(k0 + t, n - 1);
    double2 zk = {0.0, 0.0};
    int hv = 0;
    double hk[kHistDepth] = {}, solk = 0.0;
    if (t < kCircRows) {
        zk = Z[off + kt];
        if (gm_on) {
            hv = gm.stage > 0 ? gm.hvalid[gm.stage - 1] : 0;
            if (hv > 0) {
                const double* h0 = gm.hist + (long long)(gm.stage - 1) * kHistDepth * gm.total + off + kt;
#pragma unroll
                for (int l = 0; l < kHistDepth; ++l) hk[l] = l < hv ? h0[(long long)l * gm.total] : 0.0;
            } else {
                solk = gm.sol[off + kt];
            }
        }
    }
    double2 zp0 = {0.0, 0.0}, zp1 = {0.0, 0.0}, zq0 = {0.0, 0.0}, zq1 = {0.0, 0.0};
    double pp0 = 0.0, pp1 = 0.0;
    for (int j0 = 0; j0 < n; j0 += kCircChunk) {
        const int jn = min(kCircChunk, n - j0);
        __syncthreads();   // the previous chunk has been consumed
        {   // every load of the chunk in flight before the first LDS store
            double2 zv[kCircPer];
            double pv[kCircPer], ar[kCircWinPer], ai[kCircWinPer], bv[kCircWinPer];
#pragma unroll
            for (int e = 0; e < kCircPer; ++e) {
                const int c = t + 256 * e;
                zv[e] = c < jn ? Z[off + j0 + c] : make_double2(0.0, 0.0);
                pv[e] = c < jn ? phi[off + j0 + c] : 0.0;
            }
            int q[kCircWinPer];
            circ_window(k0, j0, n, q);
#pragma unroll
            for (int e = 0; e < kCircWinPer; ++e) {
                ar[e] = c1r[q[e]];
                ai[e] = c1i[q[e]];
                bv[e] = c2[q[e]];
            }
#pragma unroll
            for (int e = 0; e < kCircPer; ++e) {
                const int c = t + 256 * e;
                if (c < jn) {
                    const int j = j0 + c;
                    L.z[c] = make_double2(zv[e].x - 2 * kPi * (double)j / n, zv[e].y);
                    L.p[c] = pv[e] - plin * (double)j;
                }
            }
#pragma unroll
            for (int e = 0; e < kCircWinPer; ++e) {
                const int u = t + 256 * e;
                if (u < kCircWin) {
                    L.c1[u] = make_double2(ar[e], ai[e]);
                    L.c2[u] = bv[e];
                }
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int c = lane; c < jn; c += 64) {
            const double2 z = L.z[c];
            const double p = L.p[c];
            const int u0 = w - c + kCircChunk - 1;
            const double2 a0 = L.c1[u0], a1 = L.c1[u0 + 4];
            const double b0 = L.c2[u0], b1 = L.c2[u0 + 4];
            zp0.x += a0.x * z.x - a0.y * z.y;
            zp0.y += a0.x * z.y + a0.y * z.x;
            zp1.x += a1.x * z.x - a1.y * z.y;
            zp1.y += a1.x * z.y + a1.y * z.x;
            zq0.x += b0 * z.x;
            zq0.y += b0 * z.y;
            zq1.x += b1 * z.x;
            zq1.y += b1 * z.y;
            pp0 += a0.x * p;
            pp1 += a1.x * p;
        }
    }
    zp0.x = wave_sum(zp0.x); zp0.y = wave_sum(zp0.y);
    zp1.x = wave_sum(zp1.x); zp1.y = wave_sum(zp1.y);
    zq0.x = wave_sum(zq0.x); zq0.y = wave_sum(zq0.y);
    zq1.x = wave_sum(zq1.x); zq1.y = wave_sum(zq1.y);
    pp0 = wave_sum(pp0);
    pp1 = wave_sum(pp1);
    if (lane < 2) {   // lane 0: row w, lane 1: row w + 4
        const int k = k0 + w + 4 * lane;
        if (k < n) {
            const double2 zp = lane ? zp1 : zp0, zq = lane ? zq1 : zq0;
            const double pp = lane ? pp1 : pp0;
            Zp[off + k] = make_double2(zp.x * xs + xs, zp.y * xs);                 // :323,374
            const double s2 = 4.0 * kPi * kPi / ((double)n * n);                  // :321
            Zpp[off + k] = make_double2(zq.x * s2, zq.y * s2);
            double r = pp * xs;                                                    // :324, :299
            if (ph.U != 0) r += -(1 + ph.rho) * kPi * ph.U / n;                   // :380-383
            b[off + k] = r;
            work[off + k] = r;
        }
    }
    if (t < kCircRows && k0 + t < n) {
        const int k = k0 + t;
        if (trig) trig[off + k] = node_trig(zk);
        // GMRES initial guess of this RK stage: its mu over the previous steps extrapolated
        // (error O(dt^5) once five are kept), else the last solve's mu
        if (gm_on) gm.x[off + k] = hv > 0 ? extrapolate_values(hk, hv) : solk;
    }
}

// mu' = D1 mu * 2 pi / n (BaseBoundaryIntegrator.cuh:201-203), mu real, for kCircRows rows of one
// surface.  grid = (ceil(max n / kCircRows), surfaces).
__global__ __launch_bounds__(256) void mu_circ_kernel(const SurfDesc* __restrict__ sd,
                                                      const double* __restrict__ circ,
                                                      const double* __restrict__ mu,
                                                      double2* __restrict__ mup) {
    __shared__ double m[kCircChunk];
    __shared__ double2 c1[kCircWin];
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int k0 = blockIdx.x * kCircRows;
    if (k0 >= n) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const double* c1r = circ + d.coff;
    const double* c1i = c1r + n;
    double2 a0s = {0.0, 0.0}, a1s = {0.0, 0.0};
    for (int j0 = 0; j0 < n; j0 += kCircChunk) {
        const int jn = min(kCircChunk, n - j0);
        __syncthreads();
        {   // every load of the chunk in flight before the first LDS store
            double mv[kCircPer], ar[kCircWinPer], ai[kCircWinPer];
#pragma unroll
            for (int e = 0; e < kCircPer; ++e) {
                const int c = t + 256 * e;
                mv[e] = c < jn ? mu[off + j0 + c] : 0.0;
            }
            int q[kCircWinPer];
            circ_window(k0, j0, n, q);
#pragma unroll
            for (int e = 0; e < kCircWinPer; ++e) {
                ar[e] = c1r[q[e]];
                ai[e] = c1i[q[e]];
            }
#pragma unroll
            for (int e = 0; e < kCircPer; ++e) {
                const int c = t + 256 * e;
                if (c < jn) m[c] = mv[e];
            }
#pragma unroll
            for (int e = 0; e < kCircWinPer; ++e) {
                const int u = t + 256 * e;
                if (u < kCircWin) c1[u] = make_double2(ar[e], ai[e]);
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int c = lane; c < jn; c += 64) {
            const double x = m[c];
            const int u0 = w - c + kCircChunk - 1;
            const double2 a0 = c1[u0], a1 = c1[u0 + 4];
            a0s.x += a0.x * x;
            a0s.y += a0.y * x;
            a1s.x += a1.x * x;
            a1s.y += a1.y * x;
        }
    }
    a0s.x = wave_sum(a0s.x); a0s.y = wave_sum(a0s.y);
    a1s.x = wave_sum(a1s.x); a1s.y = wave_sum(a1s.y);
    if (lane < 2) {
        const int k = k0 + w + 4 * lane;
        if (k < n) {
            const double2 a = lane ? a1s : a0s;
            const double xs = 2.0 * kPi / n;
            mup[off + k] = make_double2(a.x * xs, a.y * xs);
        }
    }
}

}  // namespace sfh
