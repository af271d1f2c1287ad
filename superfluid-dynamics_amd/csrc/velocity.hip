// velocity.hip -- surface velocity, Bernoulli right-hand side and RK4 stage update, fused.
//
// Replaces createVelocityMatrices / createHeliumVelocityMatrices (CuSuperHelium/
// WaterVelocities.cuh:38-107), which materialise the N x N complex V1 (64 MB at N = 2048,
// written twice per RHS), the Zgemv pair of VelocityCalculator::calculateVelocities (:205-242),
// conjugate_vector, compute_rhs_phi_expression / compute_rhs_helium_phi_expression
// (createM.cuh:96-117,171-211) and the Zaxpy / add_k_vectors stages of
// AutonomousRungeKuttaStepper::runStep (AutonomousRungeKuttaStepper.cuh:123-361).
//
// One wave per node k.  The lanes sweep j != k accumulating
//   S_k = sum_j mu_j cot((z_k - z_j)/2)      Si_k = sum_j mu_j cot((z_k - conj z_j)/2 + i h)
// and lane 0 finishes
//   w_k = V2_k mu'_k + V1_kk mu_k + i (-1/4pi) S_k [+ i (1/4pi) Si_k]
//   V2_k = i / (2 pi Z'_k),  V1_kk = i (-1/4pi) Z''_k / Z'_k^2 [+ i (1/4pi) cot(i (y_k + h))] +- 1/(2 Z'_k)
//   q_k = conj(w_k),  dphi_k = Bernoulli(q_k, y_k),  then the RK4 stage of node k.
#include "sfh_kernels.hpp"
#include "sfh_physics.hpp"

namespace sfh {

// Lower (q1) and upper (q2) fluid velocity at node k of one surface; every lane of the wave
// takes part and every lane returns the result.
__device__ inline void node_velocity(const SurfDesc& d, const NodeTrig* __restrict__ tr,
                                     const double2* __restrict__ z, const double* __restrict__ m,
                                     const double2* __restrict__ Zp, const double2* __restrict__ Zpp,
                                     const double2* __restrict__ mup, const Phys& ph, int k, int lane,
                                     double2& q1, double2& q2) {
    const int n = d.n;
    const NodeTrig tk = tr[k];
    const double h = ph.depth, eh = exp(h), ehi = exp(-h);
    double sx = 0.0, sy = 0.0, ix = 0.0, iy = 0.0;
    for (int j0 = lane; j0 < n; j0 += 4 * 64) {
        NodeTrig tj[4];
        double mj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 64;
            if (j < n) {
                tj[u] = tr[j];
                mj[u] = m[j];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 64;
            if (j >= n || j == k) continue;
            const double2 ct = cot_pair(tk, tj[u]);
            sx += mj[u] * ct.x;
            sy += mj[u] * ct.y;
            if (ph.image) {
                const double2 cim = cot_image(tk, tj[u], eh, ehi);
                ix += mj[u] * cim.x;
                iy += mj[u] * cim.y;
            }
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (ph.image) {
        ix = wave_sum(ix);
        iy = wave_sum(iy);
    }
    const int g = d.voff + k;
    velocity_from_sums(ph, Zp[g], Zpp[g], mup[g], m[k], z[k].y, sx, sy, ix, iy, q1, q2);
}

__global__ __launch_bounds__(256) void velocity_stage_kernel(const SurfDesc* __restrict__ sd,
                                                             const NodeTrig* __restrict__ trig,
                                                             const double2* __restrict__ Z,
                                                             const double2* __restrict__ Zp,
                                                             const double2* __restrict__ Zpp,
                                                             const double* __restrict__ mu,
                                                             const double2* __restrict__ mup,
                                                             Phys ph, StageCtl ctl) {
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= n) return;
    const double2* z = Z + d.voff;
    const double* m = mu + d.voff;
    const NodeTrig* tr = trig + d.voff;
    double2 q1, q2;
    node_velocity(d, tr, z, m, Zp, Zpp, mup, ph, k, lane, q1, q2);
    // createM.cuh:104 reads V1[1].real() (lower velocity of node 1) in the rho * q1.q2 term;
    // with rho != 0 node 1 is recomputed here (each surface uses its own node 1).
    double2 q1q = q1;
    if (ph.kind == 0 && ph.rho_quirk && ph.rho != 0.0 && n > 1 && k != 1) {
        double2 u1, u2;
        node_velocity(d, tr, z, m, Zp, Zpp, mup, ph, 1, lane, u1, u2);
        q1q = u1;
    }
    if (lane != 0) return;
    const int g = d.voff + k;
    if (ctl.qup) ctl.qup[g] = q2;
    if (!ph.need_upper) q2 = make_double2(0.0, 0.0);
    const double dphi = bernoulli(ph, z[k].y, q1, q2, q1q, Zp[g], Zpp[g]);

    stage_update(ctl, g, q1, dphi);
}

}  // namespace sfh
