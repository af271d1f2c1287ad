// velocity.hip -- surface velocity, Bernoulli right-hand side and RK4 stage update, fused.
//
// Replaces createVelocityMatrices / createHeliumVelocityMatrices (CuSuperHelium/
// WaterVelocities.cuh:38-107), which materialise the N x N complex V1 (64 MB at N = 2048,
// written twice per RHS), the Zgemv pair of VelocityCalculator::calculateVelocities (:205-242),
// conjugate_vector, compute_rhs_phi_expression / compute_rhs_helium_phi_expression
// (createM.cuh:96-117,171-211) and the Zaxpy / add_k_vectors stages of
// AutonomousRungeKuttaStepper::runStep (AutonomousRungeKuttaStepper.cuh:123-361).
//
// A workgroup owns kPairRows nodes k.  Every thread takes the columns j = t (mod 256) (each
// workgroup starting at its own offset), loads the node's half-angle table entry and mu_j once per
// column and adds its terms to the sums of all the workgroup's rows,
//   S_k = sum_j mu_j cot((z_k - z_j)/2)      Si_k = sum_j mu_j cot((z_k - conj z_j)/2 + i h),
// the sums are reduced over the workgroup, and one thread per row finishes
//   w_k = V2_k mu'_k + V1_kk mu_k + i (-1/4pi) S_k [+ i (1/4pi) Si_k]
//   V2_k = i / (2 pi Z'_k),  V1_kk = i (-1/4pi) Z''_k / Z'_k^2 [+ i (1/4pi) cot(i (y_k + h))] +- 1/(2 Z'_k)
//   q_k = conj(w_k),  dphi_k = Bernoulli(q_k, y_k),  then the RK4 stage of node k.
#include "sfh_kernels.hpp"
#include "sfh_physics.hpp"

namespace sfh {

__global__ __launch_bounds__(256) void velocity_stage_kernel(const SurfDesc* __restrict__ sd,
                                                             const NodeTrig* __restrict__ trig,
                                                             const double2* __restrict__ Z,
                                                             const double2* __restrict__ Zp,
                                                             const double2* __restrict__ Zpp,
                                                             const double* __restrict__ mu,
                                                             const double2* __restrict__ mup,
                                                             Phys ph, StageCtl ctl) {
    constexpr int R = kPairRows;
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n;
    const int k0 = blockIdx.x * R;
    if (k0 >= n) return;
    const int t = threadIdx.x;
    const double2* z = Z + d.voff;
    const double* m = mu + d.voff;
    const NodeTrig* tr = trig + d.voff;
    // createM.cuh:104 reads V1[1].real() (lower velocity of node 1) in the rho * q1.q2 term; with
    // rho != 0 every workgroup also sums node 1 (entry R of the row arrays; each surface its own)
    const bool quirk = ph.kind == 0 && ph.rho_quirk && ph.rho != 0.0 && n > 1;
    const double h = ph.depth, eh = exp(h), ehi = exp(-h);
    NodeTrig tk[R + 1];
#pragma unroll
    for (int r = 0; r < R; ++r) tk[r] = tr[min(k0 + r, n - 1)];
    tk[R] = tr[min(1, n - 1)];
    double sx[R + 1], sy[R + 1], ix[R], iy[R];
#pragma unroll
    for (int r = 0; r <= R; ++r) sx[r] = sy[r] = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) ix[r] = iy[r] = 0.0;
    // the finishing threads' own inputs (and those of node 1 for the quirk), loaded before the
    // sweep so that their latency overlaps it
    const int kt = min(k0 + min(t, R - 1), n - 1), gt = d.voff + kt;
    double2 zp_t = {0.0, 0.0}, zpp_t = {0.0, 0.0}, mup_t = {0.0, 0.0};
    double m_t = 0.0, y_t = 0.0;
    StageIn st{};
    double2 zp_1 = {0.0, 0.0}, zpp_1 = {0.0, 0.0}, mup_1 = {0.0, 0.0};
    double m_1 = 0.0, y_1 = 0.0;
    if (t < R) {
        zp_t = Zp[gt];
        zpp_t = Zpp[gt];
        mup_t = mup[gt];
        m_t = m[kt];
        y_t = z[kt].y;
        st = stage_load(ctl, gt);
        if (quirk) {
            const int g1 = d.voff + 1;
            zp_1 = Zp[g1];
            zpp_1 = Zpp[g1];
            mup_1 = mup[g1];
            m_1 = m[1];
            y_1 = z[1].y;
        }
    }
    const int shift = (int)(((long long)blockIdx.x * 64) % n);
    for (int j0 = 0; j0 < n; j0 += 8 * 256) {
        NodeTrig tj[8];
        double mj[8];
        int jc[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {   // unconditional (clamped) loads: no branch between them
            int j = min(j0 + t + 256 * e, n - 1) + shift;
            if (j >= n) j -= n;
            jc[e] = j;
            tj[e] = tr[j];
            mj[e] = m[j];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (j0 + t + 256 * e >= n) break;
            const int j = jc[e];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (k0 + r >= n) break;
                if (j == k0 + r) continue;
                const CotParts cp = cot_parts(tk[r], tj[e]);
                const double w = mj[e] * cp.id;
                sx[r] += w * cp.p;
                sy[r] -= w * cp.q;
                if (ph.image) {
                    const CotParts cq = cot_image_parts(tk[r], tj[e], eh, ehi);
                    const double wi = mj[e] * cq.id;
                    ix[r] += wi * cq.p;
                    iy[r] -= wi * cq.q;
                }
            }
            if (quirk && j != 1) {
                const CotParts cp = cot_parts(tk[R], tj[e]);
                const double w = mj[e] * cp.id;
                sx[R] += w * cp.p;
                sy[R] -= w * cp.q;
            }
        }
    }
    __shared__ double red[2 * R * 256];
    __shared__ double so[2 * R], io[2 * R], qo[2];
    {
        double v[2 * R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = sx[r];
            v[R + r] = sy[r];
        }
        rows_reduce<2 * R>(v, red, so);
    }
    if (ph.image) {
        double v[2 * R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = ix[r];
            v[R + r] = iy[r];
        }
        rows_reduce<2 * R>(v, red, io);
    }
    if (quirk) {
        const double v[2] = {sx[R], sy[R]};
        rows_reduce<2>(v, red, qo);
    }
    if (t >= R || k0 + t >= n) return;
    const int k = k0 + t, g = d.voff + k;
    double2 q1, q2;
    velocity_from_sums(ph, zp_t, zpp_t, mup_t, m_t, y_t, so[t], so[R + t], ph.image ? io[t] : 0.0,
                       ph.image ? io[R + t] : 0.0, q1, q2);
    double2 q1q = q1;
    if (quirk && k != 1) {
        double2 u2;
        velocity_from_sums(ph, zp_1, zpp_1, mup_1, m_1, y_1, qo[0], qo[1], 0.0, 0.0, q1q, u2);
    }
    if (ctl.qup) ctl.qup[g] = q2;
    if (!ph.need_upper) q2 = make_double2(0.0, 0.0);
    const double dphi = bernoulli(ph, y_t, q1, q2, q1q, zp_t, zpp_t);
    stage_store(ctl, g, q1, dphi, st);
}

}  // namespace sfh
