// deriv.hip -- spectral derivative of a standalone field by in-LDS FFT, one workgroup per surface.
//
// Replaces FftDerivative::exec (CuSuperHelium/Derivatives.cuh:189-257) for the
// calculateDerivativeFFT256 entry point (the right-hand side itself differentiates with the
// circulant products of spectral.hip).  The workgroup loads the field into LDS in digit-reversed order,
// transforms it forward in place (radix-2 decimation in time, plus a radix-3 combine when
// n = 3 * 2^m), applies the reference multiplier table (utilities.cuh:106-148 first derivative,
// :178-200 second), transforms back with the mirror-image decimation-in-frequency pass (so the
// result lands in the same digit-reversed slots and no permutation pass is needed).
#include "sfh_kernels.hpp"
#include "sfh_fft.hpp"
#include "solver.hpp"

namespace sfh {

// forward transform, multiplier, inverse transform; leaves sample j at s[lds_index(j)].
__device__ inline void derivative_in_lds(double2* s, const SurfDesc& d, const double2* tw, bool second) {
    fft_dit(s, d, -1, tw);
    for (int i = threadIdx.x; i < d.n; i += blockDim.x)
        s[i] = second ? second_mult(s[i], i, d.n) : first_mult(s[i], i, d.n);
    __syncthreads();
    fft_dif(s, d, +1, tw);
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

namespace sfh {

// createInitialBatchedZ (JacobianCalculator.cuh:11-76) for +eps: batch b = c n + p perturbs
// coordinate c (0: x, 1: y, 2: phi) of node p.  out (complex, 6 n^2): the positions of batch b at
// [b n, b n + n), its potentials phi + 0i at [3 n^2 + b n, 3 n^2 + b n + n).
__global__ __launch_bounds__(256) void perturbed_states_kernel(const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ phi, double2* __restrict__ out,
                                                               int n, double eps) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (j >= n) return;
    const int c = b / n, p = b % n;
    double2 z = make_double2(x[j], y[j]);
    double ph = phi[j];
    if (j == p) {
        if (c == 0) z.x += eps;
        else if (c == 1) z.y += eps;
        else ph += eps;
    }
    out[(long long)b * n + j] = z;
    out[3LL * n * n + (long long)b * n + j] = make_double2(ph, 0.0);
}

void perturbed_states(const double* x, const double* y, const double* phi, int n, double eps, double* out_host) {
    if (n < 1 || n > 4096) throw ArgError("unsupported node count " + std::to_string(n));
    double *dx = nullptr, *dout = nullptr;
    const size_t outn = 12ull * n * n;   // doubles
    SFH_HIP(hipMalloc(&dx, 3 * (size_t)n * sizeof(double)));
    hipError_t e = hipMalloc(&dout, outn * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(dx);
        SFH_HIP(e);
    }
    try {
        SFH_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
        SFH_HIP(hipMemcpy(dx + n, y, n * sizeof(double), hipMemcpyHostToDevice));
        SFH_HIP(hipMemcpy(dx + 2 * n, phi, n * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(perturbed_states_kernel, dim3((n + 255) / 256, 3 * n), dim3(256), 0, 0, dx, dx + n,
                           dx + 2 * n, reinterpret_cast<double2*>(dout), n, eps);
        SFH_HIP(hipGetLastError());
        SFH_HIP(hipMemcpy(out_host, dout, outn * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        (void)hipFree(dx);
        (void)hipFree(dout);
        throw;
    }
    SFH_HIP(hipFree(dx));
    SFH_HIP(hipFree(dout));
}

}  // namespace sfh
