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
