// assemble.hip -- dipole matrix M = I/2 + K in FP64, one wave per row.
//
// Replaces createMKernel (CuSuperHelium/createM.cuh:43-63) and createFiniteDepthMKernel
// (:65-92), which run one thread per entry on 16x16 blocks.  Here a workgroup owns kPairRows
// rows: every thread takes the columns j = t (mod 256), loads the node's half-angle table entry
// (NodeTrig) once and forms the entries of all the workgroup's rows with it, so each table entry
// is read once per kPairRows rows, a thread has kPairRows independent entries in flight per
// column, and a wave's stores to one row are 512 contiguous bytes.  cot comes from the table with
// no transcendental call per entry.
//   M_kk = (1+rho)/2 + (1-rho)/(4 pi) Im(Z''_k / Z'_k)                       (water)
//   M_kj = (1-rho)/(4 pi) Im(Z'_k cot((z_k - z_j)/2))
//   helium: rho -> 0 and, with finite depth h, the image terms
//   M_kk -= Im(cot(i (y_k + h)))/(4 pi),  M_kj -= Im(cot((z_k - conj z_j)/2 + i h))/(4 pi)
//
// With the GMRES solver the kernel also starts the solve: while it writes row k it accumulates
// (M x0)_k and stores r0_k = b_k - (M x0)_k as the first (unnormalised) basis vector; the first
// Arnoldi step normalises it.
#include "sfh_kernels.hpp"

namespace sfh {

// The rows of M (stored as T: FP64, or FP32 rounded from the FP64 entries; not stored when M is
// null) and, with gm_on, r = b - M x in FP64 for x = the initial guess, or (residual pass) x = the
// solution, whose own rows the pass then copies into the initial guess of the next solve.
template <typename T, int R>
__device__ inline void assemble_rows(const SurfDesc* __restrict__ sd, const NodeTrig* __restrict__ trig,
                                     const double2* __restrict__ Z, const double2* __restrict__ Zp,
                                     const double2* __restrict__ Zpp, T* __restrict__ M, const Phys& ph,
                                     const int* __restrict__ only_if, const GmArgs& gm, int gm_on, bool residual) {
    if (only_if && only_if[2 * blockIdx.y] == 0) return;   // per-surface status word
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n;
    const int k0 = blockIdx.x * R;
    if (k0 >= n) return;
    const int t = threadIdx.x;
    if (residual && blockIdx.x == 0 && t == 0) reinterpret_cast<int*>(gm.ctl)[kGmCtlInts * blockIdx.y] = 0;   // done
    const NodeTrig* tr = trig + d.voff;
    const double* x0 = gm_on ? (residual ? gm.sol : gm.x) + d.voff : nullptr;
    const bool water = ph.kind == 0;
    const double c = water ? 0.25 * (1 - ph.rho) / kPi : 0.25 / kPi;
    const double ci = 0.25 / kPi;
    const double h = ph.depth, eh = exp(h), ehi = exp(-h);
    // the rows' constants (the same in every thread)
    NodeTrig tk[R];
    double czx[R], czy[R];   // c Z'_k
    double diag[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = min(k0 + r, n - 1);
        tk[r] = tr[k];
        const double2 zpk = Zp[d.voff + k];
        czx[r] = c * zpk.x;
        czy[r] = c * zpk.y;
        const double2 q = cdiv(Zpp[d.voff + k], zpk);
        diag[r] = (water ? 0.5 * (1 + ph.rho) : 0.5) + c * q.y;
        if (ph.image) diag[r] -= ci * cot64(0.0, Z[d.voff + k].y + h).y;
    }
    // b of the rows whose residual the threads t < R store at the end
    const double bt = (gm_on && t < R) ? gm.b[d.voff + min(k0 + t, n - 1)] : 0.0;
    T* Mk = M ? M + d.moff + (long long)k0 * n : nullptr;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // columns in passes of 8 per thread: a pass loads the table entries of its 8 columns (all in
    // flight) and then forms their entries of every row.  Each workgroup starts its sweep at its
    // own column offset (a multiple of 64): rows are n * 8 bytes apart, so workgroups sweeping in
    // lockstep from column 0 would all store to the same few memory channels at once.
    const int shift = (int)(((long long)blockIdx.x * 64) % n);
    for (int j0 = 0; j0 < n; j0 += 8 * 256) {
        NodeTrig tj[8];
        double xj[8];
        int jc[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {   // unconditional (clamped) loads: no branch between them
            int j = min(j0 + t + 256 * e, n - 1) + shift;
            if (j >= n) j -= n;
            jc[e] = j;
            tj[e] = tr[j];
            xj[e] = gm_on ? x0[j] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (j0 + t + 256 * e >= n) break;
            const int j = jc[e];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int k = k0 + r;
                if (k >= n) break;
                double v;
                if (j == k) {
                    v = diag[r];
                } else {
                    // c Im(Z'_k cot) = id (c y'_k p - c x'_k q)
                    const CotParts cp = cot_parts(tk[r], tj[e]);
                    v = cp.id * (czy[r] * cp.p - czx[r] * cp.q);
                    if (ph.image) {
                        const CotParts ci2 = cot_image_parts(tk[r], tj[e], eh, ehi);
                        v += ci * (ci2.q * ci2.id);
                    }
                }
                if (Mk) Mk[(long long)r * n + j] = static_cast<T>(v);
                acc[r] += v * xj[e];
            }
        }
    }
    if (!gm_on) return;
    __shared__ double red[(R < 4 ? 4 : R) * 256];
    __shared__ double out[R];
    rows_reduce<R>(acc, red, out);
    if (t < R && k0 + t < n) {
        gm.V[(long long)(gm.m + 1) * d.voff + k0 + t] = bt - out[t];
        if (residual) gm.x[d.voff + k0 + t] = x0[k0 + t];
    }
}

template <int R>
__global__ __launch_bounds__(256) void assemble64_kernel(const SurfDesc* __restrict__ sd,
                                                         const NodeTrig* __restrict__ trig,
                                                         const double2* __restrict__ Z,
                                                         const double2* __restrict__ Zp,
                                                         const double2* __restrict__ Zpp,
                                                         double* __restrict__ M, Phys ph,
                                                         const int* __restrict__ only_if, GmArgs gm, int gm_on) {
    assemble_rows<double, R>(sd, trig, Z, Zp, Zpp, M, ph, only_if, gm, gm_on, false);
}

// Mixed precision: M stored in FP32, r0 = b - M x0 from the FP64 entries.
template <int R>
__global__ __launch_bounds__(256) void assemble32_kernel(const SurfDesc* __restrict__ sd,
                                                         const NodeTrig* __restrict__ trig,
                                                         const double2* __restrict__ Z,
                                                         const double2* __restrict__ Zp,
                                                         const double2* __restrict__ Zpp,
                                                         float* __restrict__ M, Phys ph, GmArgs gm) {
    assemble_rows<float, R>(sd, trig, Z, Zp, Zpp, M, ph, nullptr, gm, 1, false);
}

// Mixed precision, refinement: r = b - M mu in FP64, matrix-free (the assembly's entries, not
// stored), into the first basis vector; mu becomes the next solve's initial guess, and the
// control block is re-armed (done = 0) for that solve.
template <int R>
__global__ __launch_bounds__(256) void residual64_kernel(const SurfDesc* __restrict__ sd,
                                                         const NodeTrig* __restrict__ trig,
                                                         const double2* __restrict__ Z,
                                                         const double2* __restrict__ Zp,
                                                         const double2* __restrict__ Zpp, Phys ph, GmArgs gm) {
    assemble_rows<double, R>(sd, trig, Z, Zp, Zpp, static_cast<double*>(nullptr), ph, nullptr, gm, 1, true);
}

// rows per workgroup the solver may launch with (chosen per surface size by the autotuner)
#define SFH_ASSEMBLE_INSTANCES(R)                                                                             \
    template __global__ void assemble64_kernel<R>(const SurfDesc*, const NodeTrig*, const double2*, const double2*, \
                                                  const double2*, double*, Phys, const int*, GmArgs, int);        \
    template __global__ void assemble32_kernel<R>(const SurfDesc*, const NodeTrig*, const double2*, const double2*, \
                                                  const double2*, float*, Phys, GmArgs);                          \
    template __global__ void residual64_kernel<R>(const SurfDesc*, const NodeTrig*, const double2*, const double2*, \
                                                  const double2*, Phys, GmArgs);
SFH_ASSEMBLE_INSTANCES(1)
SFH_ASSEMBLE_INSTANCES(2)
SFH_ASSEMBLE_INSTANCES(4)

}  // namespace sfh
