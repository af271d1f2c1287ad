// deriv.hip -- spectral tangential derivatives, one workgroup per (surface, output field).
//
// Replaces ZPhiDerivative::exec (CuSuperHelium/Derivatives.cuh:310-384) and
// FftDerivative::exec (:189-257), which run 3 cuFFT forward/inverse pairs plus ~10 small
// kernels per RHS.  Each workgroup loads one field into LDS in digit-reversed order,
// transforms it forward in place (radix-2 decimation in time, plus a radix-3 combine when
// n = 3 * 2^m), applies the reference multiplier table (utilities.cuh:106-148 first derivative,
// :178-200 second), transforms back with the mirror-image decimation-in-frequency pass (so the
// result lands in the same digit-reversed slots and no permutation pass is needed) and adds the
// linear parts on the way out.
#include "sfh_gmres.hpp"
#include "sfh_fft.hpp"

namespace sfh {

// forward transform, multiplier, inverse transform; leaves sample j at s[lds_index(j)].
__device__ inline void derivative_in_lds(double2* s, const SurfDesc& d, const double2* tw, bool second) {
    fft_dit(s, d, -1, tw);
    for (int i = threadIdx.x; i < d.n; i += blockDim.x)
        s[i] = second ? second_mult(s[i], i, d.n) : first_mult(s[i], i, d.n);
    __syncthreads();
    fft_dif(s, d, +1, tw);
}

// blockIdx.x = surface, blockIdx.y = field: 0 -> Z', 1 -> Z'', 2 -> Re Phi' (into b and work).
__global__ __launch_bounds__(1024) void zphi_deriv_kernel(const SurfDesc* __restrict__ sd,
                                                          const double2* __restrict__ tw_all,
                                                          const double2* __restrict__ Z,
                                                          const double* __restrict__ phi,
                                                          double2* __restrict__ Zp,
                                                          double2* __restrict__ Zpp,
                                                          double* __restrict__ b,
                                                          double* __restrict__ work, Phys ph,
                                                          int* __restrict__ status, int lds_tw,
                                                          NodeTrig* __restrict__ trig, GmArgs gm, int gm_on) {
    extern __shared__ double2 s[];
    const SurfDesc d = sd[blockIdx.x];
    const int field = blockIdx.y;
    // re-arm the surface's interchange flag for the solve that follows in stream order
    if (field == 0 && threadIdx.x == 0 && status) status[2 * blockIdx.x] = 0;
    const int n = d.n, off = d.voff;
    const double2* tw = stage_twiddles(s, d, tw_all, lds_tw);
    const double plin = -(1 + ph.rho) * kPi * ph.U / n;   // Derivatives.cuh:286
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 v;
        if (field < 2) {
            v = Z[off + j];
            if (field == 0 && trig) trig[off + j] = node_trig(v);
            v.x -= 2 * kPi * (double)j / n;                 // Derivatives.cuh:285,317
        } else {
            v = make_double2(phi[off + j] - plin * (double)j, 0.0);
        }
        s[lds_index(d, j)] = v;
    }
    __syncthreads();
    derivative_in_lds(s, d, tw, field == 1);
    double bb = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double2 v = s[lds_index(d, j)];
        if (field == 0) {
            const double sc = 2.0 * kPi / n;
            Zp[off + j] = make_double2(v.x * sc + 2.0 * kPi / n, v.y * sc);   // :323,374
        } else if (field == 1) {
            const double sc = 4.0 * kPi * kPi / (n * n);                      // :321
            Zpp[off + j] = make_double2(v.x * sc, v.y * sc);
        } else {
            double r = v.x * (2.0 * kPi / n);                                  // :324, :299
            if (ph.U != 0) r += -(1 + ph.rho) * kPi * ph.U / n;               // :380-383
            b[off + j] = r;
            work[off + j] = r;
            bb += r * r;
        }
    }
    // GMRES: |b| for the stopping test, and the solve's control block re-armed (the Arnoldi steps
    // that follow in stream order read it)
    if (gm_on && field == 2) {
        __shared__ double red[kDerivThreads / 64];
        bb = wave_sum(bb);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bb;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
            GmCtl* c = reinterpret_cast<GmCtl*>(gm.ctl) + blockIdx.x;
            c->bnorm = sqrt(t);
            c->done = 0;
            c->iters = 0;
        }
    }
    // GMRES initial guess of this RK stage: its mu over the previous steps extrapolated (error
    // O(dt^3) once three are kept), else the last solve's mu (kept in place)
    if (gm_on && field == 2 && gm.stage > 0) {
        const int hv = gm.hvalid[gm.stage - 1];
        if (hv > 0) {
            const double* h0 = gm.hist + (long long)(gm.stage - 1) * kHistDepth * gm.total + off;
            for (int j = threadIdx.x; j < n; j += blockDim.x) gm.x[off + j] = extrapolate_history(h0 + j, gm.total, hv);
        }
    }
}

// mu' = IFFT(first_mult(FFT(mu))) * 2 pi / n, mu real (BaseBoundaryIntegrator.cuh:201-203).
// With the GMRES solver it first finishes the solve of its surface: y from the rotated
// Hessenberg system R y = g (one thread), mu += sum_i y_i v_i, and the stage's history shift.
__global__ __launch_bounds__(1024) void mu_deriv_kernel(const SurfDesc* __restrict__ sd,
                                                        const double2* __restrict__ tw_all,
                                                        double* __restrict__ mu,
                                                        double2* __restrict__ mup, int lds_tw, GmArgs gm,
                                                        int gm_on) {
    extern __shared__ double2 s[];
    __shared__ double ys[kMaxKrylov];
    __shared__ int kk;
    const SurfDesc d = sd[blockIdx.x];
    const int n = d.n, off = d.voff;
    const double2* tw = stage_twiddles(s, d, tw_all, lds_tw);
    if (gm_on) {
        const GmView g = gm_view(gm, sd, blockIdx.x);
        const int m = gm.m;
        if (threadIdx.x == 0) {
            const int k = g.ctl->iters;
            for (int i = k - 1; i >= 0; --i) {   // ys holds y first, then y_i sc[i]
                double acc = g.g[i];
                for (int l = i + 1; l < k; ++l) acc -= g.H[(long long)l * (m + 1) + i] * ys[l];
                ys[i] = acc / g.H[(long long)i * (m + 1) + i];
            }
            for (int i = 0; i < k; ++i) ys[i] *= g.sc[i];
            kk = k;
        }
        __syncthreads();
        const int k = kk;
        double* h0 = gm.stage > 0 ? gm.hist + (long long)(gm.stage - 1) * kHistDepth * gm.total + off : nullptr;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            double x = mu[off + j];
            for (int i0 = 0; i0 < k; i0 += 8) {   // 8 basis loads in flight per pass
                double vv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) vv[u] = i0 + u < k ? g.V[(long long)(i0 + u) * n + j] : 0.0;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (i0 + u < k) x += ys[i0 + u] * vv[u];
            }
            mu[off + j] = x;
            if (h0) push_history(h0 + j, gm.total, x);
            s[lds_index(d, j)] = make_double2(x, 0.0);
        }
        if (h0 && blockIdx.x == 0 && threadIdx.x == 0)
            gm.hvalid[gm.stage - 1] = min(gm.hvalid[gm.stage - 1] + 1, kHistDepth);
    } else {
        for (int j = threadIdx.x; j < n; j += blockDim.x) s[lds_index(d, j)] = make_double2(mu[off + j], 0.0);
    }
    __syncthreads();
    derivative_in_lds(s, d, tw, false);
    const double sc = 2.0 * kPi / n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double2 v = s[lds_index(d, j)];
        mup[off + j] = make_double2(v.x * sc, v.y * sc);
    }
}

// Standalone derivative of an arbitrary complex field (FftDerivative::exec with scale 1 and
// no linear part), used by the calculateDerivativeFFT256 entry point.
__global__ __launch_bounds__(1024) void plain_deriv_kernel(const SurfDesc* __restrict__ sd,
                                                           const double2* __restrict__ tw_all,
                                                           const double2* __restrict__ in,
                                                           double2* __restrict__ out, int second,
                                                           double scale, int lds_tw) {
    extern __shared__ double2 s[];
    const SurfDesc d = sd[blockIdx.x];
    const int n = d.n, off = d.voff;
    const double2* tw = stage_twiddles(s, d, tw_all, lds_tw);
    for (int j = threadIdx.x; j < n; j += blockDim.x) s[lds_index(d, j)] = in[off + j];
    __syncthreads();
    derivative_in_lds(s, d, tw, second != 0);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double2 v = s[lds_index(d, j)];
        out[off + j] = (scale != 1.0) ? make_double2(v.x * scale, v.y * scale) : v;
    }
}

}  // namespace sfh
