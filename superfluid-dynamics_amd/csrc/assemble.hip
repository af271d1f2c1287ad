// assemble.hip -- dipole matrix M = I/2 + K in FP64, one wave per row.
//
// Replaces createMKernel (CuSuperHelium/createM.cuh:43-63) and createFiniteDepthMKernel
// (:65-92), which run one thread per entry on 16x16 blocks.  Here wave w of a workgroup owns
// row k; its 64 lanes sweep the columns, so Z_k, Z'_k stay in registers, Z_j loads and M row
// stores are fully coalesced, and cot comes from the per-node half-angle table (NodeTrig) with no
// transcendental call per entry.
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

__global__ __launch_bounds__(256) void assemble64_kernel(const SurfDesc* __restrict__ sd,
                                                         const NodeTrig* __restrict__ trig,
                                                         const double2* __restrict__ Z,
                                                         const double2* __restrict__ Zp,
                                                         const double2* __restrict__ Zpp,
                                                         double* __restrict__ M, Phys ph,
                                                         const int* __restrict__ only_if, GmArgs gm, int gm_on) {
    if (only_if && only_if[2 * blockIdx.y] == 0) return;   // per-surface status word
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= n) return;
    const NodeTrig* tr = trig + d.voff;
    const NodeTrig tk = tr[k];
    const double2 zpk = Zp[d.voff + k];
    const double* x0 = gm_on ? gm.x + d.voff : nullptr;
    const bool water = ph.kind == 0;
    const double c = water ? 0.25 * (1 - ph.rho) / kPi : 0.25 / kPi;
    const double ci = 0.25 / kPi;
    const double h = ph.depth, eh = exp(h), ehi = exp(-h);
    double* row = M + d.moff + (long long)k * n;
    double acc = 0.0;
    for (int j0 = lane; j0 < n; j0 += 4 * 64) {
        NodeTrig tj[4];
        double xj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 64;
            if (j < n) {
                tj[u] = tr[j];
                xj[u] = gm_on ? x0[j] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 64;
            if (j >= n) continue;
            double v;
            if (j == k) {
                const double2 r = cdiv(Zpp[d.voff + k], zpk);
                v = (water ? 0.5 * (1 + ph.rho) : 0.5) + c * r.y;
                if (ph.image) v -= ci * cot64(0.0, Z[d.voff + k].y + h).y;
            } else {
                const double2 ct = cot_pair(tk, tj[u]);
                v = c * (zpk.x * ct.y + zpk.y * ct.x);
                if (ph.image) v -= ci * cot_image(tk, tj[u], eh, ehi).y;
            }
            row[j] = v;
            acc += v * xj[u];
        }
    }
    if (!gm_on) return;
    acc = wave_sum(acc);
    if (lane == 0) gm.V[(long long)(gm.m + 1) * d.voff + k] = gm.b[d.voff + k] - acc;
}

}  // namespace sfh
