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
//
// velocity_gm1_kernel is the same kernel at the end of a GMRES solve of one Arnoldi step (the
// launch sequence of two: gm_iter_kernel's product, then this).  With v0 = r0/|r0| and
// w = M v0, the least-squares step gives mu = x0 + (v0.w / |w|^2) r0; launch 0 left the partial
// sums of v0.w and |w|^2, so every workgroup reduces them and forms mu_j = x0_j + alpha r0_j as
// it sweeps the columns.  The same sweep adds the circulant product of the derivative,
// mu'_k = (2 pi / n) sum_j c1[(k - j) mod n] mu_j (mu_circ_kernel's product;
// BaseBoundaryIntegrator.cuh:201-203) for the workgroup's rows from a window of the kernel in LDS,
// so the solution and its derivative need no launch of their own.  The workgroup stores its rows
// of mu (and pushes them into the stage's warm-start history); workgroup 0 of each surface
// rebuilds |w - (v0.w) v0| for the residual estimate and closes the control block.
#include "sfh_gmres.hpp"
#include "sfh_kernels.hpp"
#include "sfh_physics.hpp"

namespace sfh {

template <bool G1, int R>
__device__ inline void velocity_rows(const SurfDesc* __restrict__ sd, const NodeTrig* __restrict__ trig,
                                     const double2* __restrict__ Z, const double2* __restrict__ Zp,
                                     const double2* __restrict__ Zpp, const double* __restrict__ mu,
                                     const double2* __restrict__ mup, double2* __restrict__ mup_out,
                                     const Phys& ph, const StageCtl& ctl,
                                     const GmArgs& gm, const double* __restrict__ circ) {
    static_assert(R <= kCircRowsPerWG, "the rows of a workgroup fit the circulant window");
    const int s = blockIdx.y;
    const SurfDesc d = sd[s];
    const int n = d.n;
    const int k0 = blockIdx.x * R;
    if (k0 >= n) return;
    const int t = threadIdx.x;
    const double2* z = Z + d.voff;
    const double* m = G1 ? nullptr : mu + d.voff;
    const NodeTrig* tr = trig + d.voff;
    // createM.cuh:104 reads V1[1].real() (lower velocity of node 1) in the rho * q1.q2 term; with
    // rho != 0 every workgroup also sums node 1 (entry R of the row arrays; each surface its own)
    const bool quirk = !G1 && ph.kind == 0 && ph.rho_quirk && ph.rho != 0.0 && n > 1;
    const double h = ph.depth, eh = exp(h), ehi = exp(-h);
    const int shift = (int)(((long long)blockIdx.x * 64) % n);
    // ---- G1: the solve's state, the partial sums of launch 0 and the first kernel window first
    GmCtl* gc = nullptr;
    const double* x0 = nullptr;
    const double* r0 = nullptr;
    const double* c1r = nullptr;
    const double* c1i = nullptr;
    int fin = 0;
    double beta = 0.0, bnorm = 0.0;
    DotLoads dl;
    const int nwg = (n + kRowsPerWG - 1) / kRowsPerWG;
    const double* part = nullptr;
    bool fastdots = false;
    double ar[kCircWinPer], ai[kCircWinPer];
    if constexpr (G1) {
        gc = reinterpret_cast<GmCtl*>(gm.ctl) + s;
        x0 = gm.x + d.voff;
        r0 = gm.V + (long long)(gm.m + 1) * d.voff;
        c1r = circ + d.coff;
        c1i = c1r + n;
        fin = gc->done;
        beta = gc->beta;
        bnorm = gc->bnorm;
        part = gm.part + (long long)s * gm.part_stride;   // parity 0: launch 0
        fastdots = dots_fast(nwg, 2);
        if (fastdots) dots_load(part, nwg, 2, dl);
        int q[kCircWinPer];
        circ_window(k0, shift, n, q);
#pragma unroll
        for (int e = 0; e < kCircWinPer; ++e) {
            ar[e] = c1r[q[e]];
            ai[e] = c1i[q[e]];
        }
    }
    NodeTrig tk[R + 1];
#pragma unroll
    for (int r = 0; r < R; ++r) tk[r] = tr[min(k0 + r, n - 1)];
    tk[R] = tr[min(1, n - 1)];
    double sx[R + 1], sy[R + 1], ix[R], iy[R], cx[R], cy[R];
#pragma unroll
    for (int r = 0; r <= R; ++r) sx[r] = sy[r] = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) ix[r] = iy[r] = cx[r] = cy[r] = 0.0;
    // the finishing threads' own inputs (and those of node 1 for the quirk), loaded before the
    // sweep so that their latency overlaps it
    const int kt = min(k0 + min(t, R - 1), n - 1), gt = d.voff + kt;
    double2 zp_t = {0.0, 0.0}, zpp_t = {0.0, 0.0}, mup_t = {0.0, 0.0};
    double m_t = 0.0, y_t = 0.0;
    StageIn st{};
    double2 zp_1 = {0.0, 0.0}, zpp_1 = {0.0, 0.0}, mup_1 = {0.0, 0.0};
    double m_1 = 0.0, y_1 = 0.0;
    // G1: the own rows' x0, r0 and the history levels that move down (kHistDepth + 1 values per
    // row), staged through LDS rather than held in registers across the sweep
    double* h0 = nullptr;
    __shared__ double own[G1 ? (kHistDepth + 1) * R : 1];
    if constexpr (G1) {
        h0 = history(gm, sd, s);
        if (t < (kHistDepth + 1) * R) {
            const int r = t % R, f = t / R;   // f = 0: x0, 1: r0, 2..: history level f - 2
            const int kr = min(k0 + r, n - 1);
            double v = 0.0;
            if (f == 0)
                v = x0[kr];
            else if (f == 1)
                v = r0[kr];
            else if (h0)
                v = h0[(long long)(f - 2) * gm.total + kr];
            own[t] = v;
        }
    }
    if (t < R) {
        zp_t = Zp[gt];
        zpp_t = Zpp[gt];
        y_t = z[kt].y;
        st = stage_load(ctl, gt);
        if constexpr (!G1) {
            mup_t = mup[gt];
            m_t = m[kt];
        }
        if (quirk) {
            const int g1 = d.voff + 1;
            zp_1 = Zp[g1];
            zpp_1 = Zpp[g1];
            mup_1 = mup[g1];
            m_1 = m[1];
            y_1 = z[1].y;
        }
    }
    __shared__ double2 c1w[G1 ? kCircWin : 1];
    __shared__ double sums[2];
    double alpha = 0.0;   // mu = x0 + alpha r0
    if constexpr (G1) {
        if (fastdots)
            dots_finish(dl, 2, sums);
        else
            reduce_partials(part, nwg, 2, sums);
#pragma unroll
        for (int e = 0; e < kCircWinPer; ++e) {
            const int u = t + 256 * e;
            if (u < kCircWin) c1w[u] = make_double2(ar[e], ai[e]);
        }
        __syncthreads();
        // a solve finished by launch 0 (|r0| within the tolerance, done = 1) keeps x0.  done = 2 is
        // this launch's own finish, stamped by workgroup 0, which a workgroup that starts after it
        // (more workgroups than fit the GPU at once) reads here: it does not count
        const double vw = sums[0], ww = sums[1];
        alpha = (fin != 1 && ww > 0.0) ? vw / ww : 0.0;
    }
    for (int j0 = 0; j0 < n; j0 += 8 * 256) {
        if (G1 && j0 > 0) {   // the kernel window of this pass (n > 2048), one entry at a time
            int jw = j0 + shift;
            if (jw >= n) jw -= n;
            __syncthreads();   // the previous window has been consumed
            for (int u = t; u < kCircWin; u += 256) {
                int q = (k0 - jw - (kCircChunk - 1) + u) % n;
                if (q < 0) q += n;
                c1w[u] = make_double2(c1r[q], c1i[q]);
            }
            __syncthreads();
        }
        NodeTrig tj[8];
        double mj[8];
        int jc[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {   // unconditional (clamped) loads: no branch between them
            int j = min(j0 + t + 256 * e, n - 1) + shift;
            if (j >= n) j -= n;
            jc[e] = j;
            tj[e] = tr[j];
            if constexpr (G1)
                mj[e] = x0[j] + alpha * r0[j];
            else
                mj[e] = m[j];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = t + 256 * e;
            if (j0 + c >= n) break;
            const int j = jc[e];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (k0 + r >= n) break;
                if constexpr (G1) {   // window entry r - c + kCircChunk - 1 = c1[(k0 + r - j) mod n]
                    const double2 p = c1w[r - c + kCircChunk - 1];
                    cx[r] += p.x * mj[e];
                    cy[r] += p.y * mj[e];
                }
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
    __shared__ double so[2 * R], io[2 * R], co[2 * R], qo[2];
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
    if constexpr (G1) {
        double v[2 * R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = cx[r];
            v[R + r] = cy[r];
        }
        rows_reduce<2 * R>(v, red, co);
    }
    if (quirk) {
        const double v[2] = {sx[R], sy[R]};
        rows_reduce<2>(v, red, qo);
    }
    if constexpr (G1) {
        if (blockIdx.x == 0) {
            // the residual estimate of the step (rebuild of |w - (v0.w) v0|, as the Arnoldi
            // launches do) and the control block; no other workgroup reads them
            const double sc0 = beta > 0.0 ? 1.0 / beta : 0.0;
            const double h00 = sums[0];
            const double* w0 = gm.w + d.voff;
            double ss = 0.0;
            if (fin == 0) {
                for (int c = t; c < n; c += 256) {
                    const double u = w0[c] - h00 * (sc0 * r0[c]);
                    ss += u * u;
                }
            }
            __shared__ double red4[4];
            const double hn = sqrt(block_sum4(ss, red4));
            if (t == 0) {
                if (fin == 0) {
                    const double rho = sqrt(h00 * h00 + hn * hn);
                    const double res = rho > 0.0 ? beta * hn / rho : 0.0;
                    gc->res = res;
                    gc->hnext = hn;
                    if (!(res <= gm.tol * bnorm || hn == 0.0)) gc->unconverged += 1;
                    set_iters(gc, gm, 1);
                    gc->done = 2;
                }
                if (h0 && s == 0) count_history(gm);
            }
        }
    }
    if (t >= R || k0 + t >= n) return;
    const int k = k0 + t, g = d.voff + k;
    if constexpr (G1) {
        m_t = own[t] + alpha * own[R + t];
        const double xs = 2.0 * kPi / n;
        mup_t = make_double2(co[t] * xs, co[R + t] * xs);
        gm.sol[g] = m_t;
        mup_out[g] = mup_t;
        if (h0) {
#pragma unroll
            for (int l = kHistDepth - 1; l > 0; --l) h0[(long long)l * gm.total + k] = own[(l + 1) * R + t];
            h0[k] = m_t;
        }
    }
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
    double dphi = bernoulli(ph, y_t, q1, q2, q1q, zp_t, zpp_t);
    // energy and drive terms (on request only: their inputs are loaded here, not held across the sweep)
    if (ctl.energy) energy_store(ph, ctl, g, z[k].x, y_t, ctl.Pin[g], zp_t, q1);
    if (ph.drive) {
        double dD;
        dphi = drive_terms(ph, ctl, g, z[k].x, y_t, q1, dphi, dD);
        if (ph.augmented) stage_store_delayed(ctl, g, dD);
    }
    stage_store(ctl, g, q1, dphi, st);
}

template <int R>
__global__ __launch_bounds__(256) void velocity_stage_kernel(const SurfDesc* __restrict__ sd,
                                                             const NodeTrig* __restrict__ trig,
                                                             const double2* __restrict__ Z,
                                                             const double2* __restrict__ Zp,
                                                             const double2* __restrict__ Zpp,
                                                             const double* __restrict__ mu,
                                                             const double2* __restrict__ mup,
                                                             Phys ph, StageCtl ctl) {
    velocity_rows<false, R>(sd, trig, Z, Zp, Zpp, mu, mup, nullptr, ph, ctl, GmArgs{}, nullptr);
}

// Finishes a GMRES solve of one Arnoldi step (launch 0 of gm_iter_kernel ran its product), forms
// mu and mu' and runs the velocity stage.  Not for the rho quirk, whose node-1 term needs mu' of a
// row outside the workgroup's.  mup receives mu' and gm.sol mu.
// Two waves per SIMD (all 512 workgroups of an N = 2048 surface resident at once): the compiler
// keeps the kernel within 256 VGPRs, spilling 2.
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void velocity_gm1_kernel(
    const SurfDesc* __restrict__ sd, const NodeTrig* __restrict__ trig, const double2* __restrict__ Z,
    const double2* __restrict__ Zp, const double2* __restrict__ Zpp, double2* __restrict__ mup, Phys ph,
    StageCtl ctl, GmArgs gm, const double* __restrict__ circ) {
    velocity_rows<true, R>(sd, trig, Z, Zp, Zpp, nullptr, nullptr, mup, ph, ctl, gm, circ);
}

// rows per workgroup the solver may launch with (chosen per surface size by the autotuner)
#define SFH_VELOCITY_INSTANCES(R)                                                                                  \
    template __global__ void velocity_stage_kernel<R>(const SurfDesc*, const NodeTrig*, const double2*,                \
                                                      const double2*, const double2*, const double*, const double2*, \
                                                      Phys, StageCtl);                                              \
    template __global__ void velocity_gm1_kernel<R>(const SurfDesc*, const NodeTrig*, const double2*, const double2*, \
                                                    const double2*, double2*, Phys, StageCtl, GmArgs, const double*);
SFH_VELOCITY_INSTANCES(1)
SFH_VELOCITY_INSTANCES(2)
SFH_VELOCITY_INSTANCES(4)

}  // namespace sfh
