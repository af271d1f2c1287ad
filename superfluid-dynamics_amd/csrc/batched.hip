// batched.hip -- persistent per-surface stage kernel for ensembles of independent surfaces
// (every N <= 1024): ONE launch per RK4 stage advances all surfaces.
//
// Workgroup b owns surface order[b]; the order is by decreasing N, so the in-order dispatch of
// workgroups schedules the largest surfaces first and the small ones fill the machine behind them.
// A workgroup runs the whole right-hand side of its surface with no inter-workgroup
// synchronisation (BaseBoundaryIntegralCalculator::runTimeStep,
// CuSuperHelium/BaseBoundaryIntegrator.cuh:137-314, for one member of the batch):
//   A  spectral derivatives in LDS: Z' and Z'' from ONE forward transform of Z - 2 pi j / N,
//      Phi' from a second (Derivatives.cuh:310-384), plus the per-node half-angle table;
//   B  the dipole matrix M = I/2 + K (createM.cuh:43-92) into HBM, column-major per surface
//      (element (k, j) at j N + k, the reference's own layout), FP64 or FP32 (FP64 entries
//      rounded), and in the same pair loop M x for the warm start x (the first residual);
//   C  the solve M mu = b, warm-started from the surface's previous mu:
//        FP64  : GMRES on the stored FP64 matrix (classical Gram-Schmidt with one
//                re-orthogonalisation, restarted cycles of kBK vectors);
//        mixed : FP64 iterative refinement  r = b - M mu  with the residual evaluated
//                matrix-free in FP64 (same pair formula as the assembly), correction solves by
//                GMRES on the stored FP32 matrix; a correction whose GMRES residual plus
//                ||M - M32||_F |d| is within the tolerance ends the refinement without another
//                FP64 residual pass;
//   D  mu' (one more transform pair);
//   E  the cot-sum velocity, Bernoulli and the RK4 stage update (WaterVelocities.cuh:38-107,
//      createM.cuh:96-117, AutonomousRungeKuttaStepper.cuh:338-375).
//
// Thread t owns node t.  Pair loops (assembly, matrix-vector products, velocity) split the
// columns over G = 1024 / N groups so small surfaces still use the whole workgroup: thread t
// handles row t % N and columns j = t / N (mod G).  The column-major matrix makes every matrix
// load of a wave one 512-byte (FP64) or 256-byte (FP32) contiguous segment, and column data
// (node table, the vector) is read from LDS as a broadcast.
#include "sfh_wg.hpp"

namespace sfh {

__global__ __launch_bounds__(1024) void batched_stage_kernel(BatchArgs a, Phys ph, StageCtl ctl) {
    __shared__ BLds L;
    const int s = a.order[blockIdx.x];
    const SurfDesc d = a.sd[s];
    const int n = d.n, off = d.voff, t = threadIdx.x;
    const bool own = t < n;
    const int G = kBT / n > 0 ? kBT / n : 1;
    const double twopi_n = 2.0 * kPi / n;

    // ---------------------------------------------------------------- A: derivatives
    const NodeVals nv = wg_derivatives(d, a.tw_all, a.Zin, a.Pin, ph, L);
    const double2 zk = nv.z, zpk = nv.zp, zppk = nv.zpp;
    const double bk = nv.b;
    if (own) {
        a.Zp[off + t] = zpk;
        a.Zpp[off + t] = zppk;
        a.b[off + t] = bk;
    }

    // warm start: this stage's mu extrapolated over the previous steps (error O(dt^5) once five
    // are kept), else the previous solve's mu
    double x = own ? a.mu[off + t] : 0.0;
    double* h0 = nullptr;
    int hv = 0;
    if (ctl.stage > 0 && a.hist) {
        hv = a.hvalid[(ctl.stage - 1) * a.nsurf + s];
        h0 = a.hist + (long long)(ctl.stage - 1) * kHistDepth * a.total + off;
        if (own && hv >= 1) x = extrapolate_history(h0 + t, a.total, hv);
    }

    // ---------------------------------------------------------------- B: assembly
    // The pair loop also forms (M x)_k for the warm start x in FP64, the first residual of the
    // solve: FP64 from the stored entries, in the order of mv_stored; mixed from the FP64 entries
    // that are rounded to FP32 for storage, in the order of mv_free.
    const bool water = ph.kind == 0;
    const PairConst pc = wg_diagonal(ph, nv, n, L);
    if (own) L.vs[t] = x;
    __syncthreads();
    double mval = 0.0;
    double ds2 = 0.0;   // mixed: squared rounding errors of the FP32 entries (this thread's)
    if (t < G * n) {
        const int k = t % n, g = t / n;
        const NodeTrig tk = L.tr[k];
        const double2 zq = L.zp[k];
        const double czx = pc.c * zq.x, czy = pc.c * zq.y;
        if (a.mixed) {
            float* col = reinterpret_cast<float*>(a.M) + d.moff + k;
            double sx = 0.0, sy = 0.0, si = 0.0;
            for (int j = g; j < n; j += G) {
                float v;
                if (j == k) {
                    v = (float)L.diag[k];
                    const double rd = L.diag[k] - (double)v;
                    ds2 += rd * rd;
                } else {
                    const NodeTrig tj = L.tr[j];
                    const double vj = L.vs[j];
                    const CotParts cp = cot_parts(tk, tj);   // c Im(Z'_k cot) = id (c y' p - c x' q)
                    const double w = vj * cp.id;
                    sx += w * cp.p;
                    sy -= w * cp.q;
                    double e = cp.id * (czy * cp.p - czx * cp.q);
                    if (pc.image) {
                        const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
                        si -= vj * (cq.q * cq.id);
                        e += pc.ci * (cq.q * cq.id);
                    }
                    v = (float)e;
                    const double rd = e - (double)v;
                    ds2 += rd * rd;
                }
                col[(size_t)j * n] = v;
            }
            mval = pc.c * (zq.x * sy + zq.y * sx) - pc.ci * si;
        } else {
            double* col = reinterpret_cast<double*>(a.M) + d.moff + k;
            double acc0 = 0.0, acc1 = 0.0;
            int u = 0;
            for (int j = g; j < n; j += G, ++u) {
                double v;
                if (j == k) {
                    v = L.diag[k];
                } else {
                    const NodeTrig tj = L.tr[j];
                    const CotParts cp = cot_parts(tk, tj);
                    v = cp.id * (czy * cp.p - czx * cp.q);
                    if (pc.image) {
                        const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
                        v += pc.ci * (cq.q * cq.id);
                    }
                }
                col[(size_t)j * n] = v;
                if (u & 1)
                    acc1 += v * L.vs[j];
                else
                    acc0 += v * L.vs[j];
            }
            mval = acc0 + acc1;
        }
    }
    L.part[t] = mval;
    __syncthreads();   // the workgroup reads back only its own columns: a barrier suffices
    double mx0 = 0.0;   // (M x)_t
    if (own) {
        for (int q = 0; q < G; ++q) mx0 += L.part[q * n + t];
        if (a.mixed) mx0 += L.diag[t] * L.vs[t];
    }
    __syncthreads();

    // ---------------------------------------------------------------- C: solve
    const double bn = sqrt(bsum1(own ? bk * bk : 0.0, L));
    const double tol_abs = a.tol * bn;
    int unconv = 0, iters = 0, sweeps = 0;
    if (!a.mixed) {
        const double* Md = reinterpret_cast<const double*>(a.M) + d.moff;
        iters = wg_gmres([&]() { return mv_stored(Md, n, G, L); }, n, bk, x, false, tol_abs, a.cycles, L, &unconv,
                         true, own ? bk - mx0 : 0.0);
    } else {
        const float* Mf = reinterpret_cast<const float*>(a.M) + d.moff;
        const double dM = sqrt(bsum1(ds2, L));   // ||M - M32||_F
        bool done = false;
        for (; sweeps < a.max_refine; ++sweeps) {
            double mx = mx0;   // sweep 0: formed by the assembly
            if (sweeps > 0) {
                if (own) L.vs[t] = x;
                __syncthreads();
                mx = mv_free(n, G, pc, L);   // every thread: it synchronises
            }
            const double r = own ? bk - mx : 0.0;
            const double rn = sqrt(bsum1(r * r, L));
            if (!(rn > tol_abs)) {
                done = true;
                break;
            }
            double dx = 0.0;
            iters += wg_gmres([&]() { return mv_stored(Mf, n, G, L); }, n, r, dx, true,
                              fmax(1e-7 * rn, 0.25 * tol_abs), 1, L, nullptr);
            x += dx;
            // the FP64 residual of the corrected x, bounded without forming it:
            //   |b - M x| <= |r - M32 dx| + |(M - M32) dx| <= (GMRES residual) + ||M - M32||_F |dx|
            // which at steady state is far below the tolerance (|dx| is the warm start's error)
            const double est = L.scal[0];   // |r - M32 dx|, the last rotated right-hand side entry
            const double dn = sqrt(bsum1(own ? dx * dx : 0.0, L));
            if (est + dM * dn <= tol_abs) {
                ++sweeps;
                done = true;
                break;
            }
        }
        if (!done) unconv = 1;
    }
    if (t == 0) {
        int* st = a.stats + 4 * s;
        st[0] = iters;
        st[1] += unconv;
        st[2] = sweeps;
        if (h0) a.hvalid[(ctl.stage - 1) * a.nsurf + s] = min(hv + 1, kHistDepth);
    }
    if (own && h0) push_history(h0 + t, a.total, x);

    // ---------------------------------------------------------------- D: mu'
    if (own) {
        a.mu[off + t] = x;
        L.fa[lds_index(d, t)] = make_double2(x, 0.0);
    }
    wg_first_derivative(d, L, false);
    double2 mupk = make_double2(0.0, 0.0);
    if (own) {
        const double2 v = L.fa[lds_index(d, t)];
        mupk = make_double2(v.x * twopi_n, v.y * twopi_n);   // BaseBoundaryIntegrator.cuh:201-203
        a.mup[off + t] = mupk;
        L.vs[t] = x;
    }
    __syncthreads();

    // ---------------------------------------------------------------- E: velocity + stage
    double* psy = reinterpret_cast<double*>(L.fa);   // 2 kBT doubles: sy | ix
    double* piy = reinterpret_cast<double*>(L.fb);
    if (t < G * n) {
        const int k = t % n, g = t / n;
        const NodeTrig tk = L.tr[k];
        double sx = 0.0, sy = 0.0, ix = 0.0, iy = 0.0;
        for (int j = g; j < n; j += G) {
            if (j == k) continue;
            const NodeTrig tj = L.tr[j];
            const double mj = L.vs[j];
            const CotParts cp = cot_parts(tk, tj);
            const double w = mj * cp.id;
            sx += w * cp.p;
            sy -= w * cp.q;
            if (pc.image) {
                const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
                const double wi = mj * cq.id;
                ix += wi * cq.p;
                iy -= wi * cq.q;
            }
        }
        L.part[t] = sx;
        psy[t] = sy;
        psy[kBT + t] = ix;
        piy[t] = iy;
    } else {
        L.part[t] = 0.0;
        psy[t] = 0.0;
        psy[kBT + t] = 0.0;
        piy[t] = 0.0;
    }
    __syncthreads();
    double2 q1 = make_double2(0.0, 0.0), q2 = make_double2(0.0, 0.0);
    if (own) {
        double sx = 0.0, sy = 0.0, ix = 0.0, iy = 0.0;
        for (int q = 0; q < G; ++q) {
            sx += L.part[q * n + t];
            sy += psy[q * n + t];
            ix += psy[kBT + q * n + t];
            iy += piy[q * n + t];
        }
        velocity_from_sums(ph, zpk, zppk, mupk, x, zk.y, sx, sy, ix, iy, q1, q2);
    }
    // createM.cuh:104 reads the lower velocity of node 1 in the rho * q1.q2 term
    double2 q1q = q1;
    if (water && ph.rho_quirk && ph.rho != 0.0 && n > 1) {
        __syncthreads();
        if (t == 1) {
            L.scal[2] = q1.x;
            L.scal[3] = q1.y;
        }
        __syncthreads();
        q1q = make_double2(L.scal[2], L.scal[3]);
    }
    if (!own) return;
    const int gi = off + t;
    if (ctl.qup) ctl.qup[gi] = q2;
    if (!ph.need_upper) q2 = make_double2(0.0, 0.0);
    double dphi = bernoulli(ph, zk.y, q1, q2, q1q, zpk, zppk);
    if (ctl.energy) energy_store(ph, ctl, gi, zk.x, zk.y, ctl.Pin[gi], zpk, q1);
    if (ph.drive) {
        double dD;
        dphi = drive_terms(ph, ctl, gi, zk.x, zk.y, q1, dphi, dD);
        if (ph.augmented) stage_store_delayed(ctl, gi, dD);
    }
    stage_update(ctl, gi, q1, dphi);
}

}  // namespace sfh
