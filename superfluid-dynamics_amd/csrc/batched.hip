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
//                ||M - M32||_F |d| (bounded by 2^-24 ||M||_F) is within the tolerance ends the
//                refinement without another FP64 residual pass;
//   D  mu' (one more transform pair);
//   E  the cot-sum velocity, Bernoulli and the RK4 stage update (WaterVelocities.cuh:38-107,
//      createM.cuh:96-117, AutonomousRungeKuttaStepper.cuh:338-375).
//
// Thread t owns node t.  The pair loops (assembly, the matrix-free FP64 product, velocity) visit
// every unordered pair of nodes once in 64 x 64 tiles and give both rows of the pair their
// contribution (sfh_sym.hpp); the matrix is stored in the tiles' skewed layout, so every store
// and every load of the stored product is one contiguous 512-byte (FP64) or 256-byte (FP32) row
// per wave, and node data is read from LDS without bank conflicts.
#include <type_traits>

#include "sfh_sym.hpp"

namespace sfh {

__global__ __launch_bounds__(1024) void batched_stage_kernel(BatchArgs a, Phys ph, StageCtl ctl) {
    __shared__ BLds L;
    const int s = a.order[blockIdx.x];
    const SurfDesc d = a.sd[s];
    const int n = d.n, off = d.voff, t = threadIdx.x;
    const bool own = t < n;
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
    // One sweep over the unordered pairs (sfh_sym.hpp) writes both entries of each pair into the
    // tile-skewed matrix (FP64, or FP64 entries rounded to FP32) and forms (M x)_k for the warm start
    // x in FP64 from the unrounded entries: the first residual of the solve.
    const bool water = ph.kind == 0;
    const PairConst pc = wg_diagonal(ph, nv, n, L);
    const int nbk = (n + kTile - 1) / kTile;
    if (t < nbk * kTile) L.vs[t] = own ? x : 0.0;   // the padding entries stay 0 for every product
    __syncthreads();
    double dsq = 0.0;   // mixed: this thread's share of ||M||_F^2
    {
        const SymGeo sg = sym_geo(n);
        const int kr = sg.I * kTile + sg.l;
        const NodeTrig tk = L.tr[kr];
        const double2 zq = L.zp[kr];
        const double2 czk = make_double2(pc.c * zq.x, pc.c * zq.y);
        const double vk = L.vs[kr];
        double* const acc[1] = {L.part};
        auto store = [&](auto* Mt) {
            using T = typename std::remove_pointer<decltype(Mt)>::type;
            // the diagonal: offset 0 of the diagonal tiles
            if (t < nbk * kTile) {
                const double dg = own ? L.diag[t] : 0.0;
                Mt[(long long)((t >> 6) * nbk + (t >> 6)) * kTileElems + (t & 63)] = (T)dg;
                dsq += dg * dg;
            }
            sym_sweep<1>(n, acc, [&](int k, int j, int J, int r, auto mirror, auto full, double (&row)[1], double (&col)[1]) {
                // tile bases are wave-uniform; the offsets within a tile are 32-bit
                T* const tij = Mt + (long long)(sg.I * nbk + J) * kTileElems;
                T* const tji = Mt + (long long)(J * nbk + sg.I) * kTileElems;
                const unsigned l = sg.l;
                double e = 0.0, em = 0.0;
                if (decltype(full)::value || (k < n && j < n)) {
                    entry_pair(tk, L.tr[j], czk, L.zp[j], pc, e, em);
                    row[0] += e * L.vs[j];
                    dsq += e * e;
                    if constexpr (decltype(mirror)::value) {
                        col[0] += em * vk;
                        dsq += em * em;
                    }
                }
                tij[(unsigned)r * kTile + l] = (T)e;
                if constexpr (decltype(mirror)::value) tji[((unsigned)(kTile - r) & (kTile - 1)) * kTile + ((l + r) & (kTile - 1))] = (T)em;
            });
        };
        if (a.mixed)
            store(reinterpret_cast<float*>(a.M) + d.moff);
        else
            store(reinterpret_cast<double*>(a.M) + d.moff);
    }
    double mx0 = 0.0;   // (M x)_t
    if (own) mx0 = sym_total(L.part, n) + L.diag[t] * L.vs[t];
    __syncthreads();

    // ---------------------------------------------------------------- C: solve
    const double bn = sqrt(bsum1(own ? bk * bk : 0.0, L));
    const double tol_abs = a.tol * bn;
    int unconv = 0, iters = 0, sweeps = 0;
    if (!a.mixed) {
        const double* Md = reinterpret_cast<const double*>(a.M) + d.moff;
        iters = wg_gmres([&]() { return mv_tiles(Md, n, L); }, n, bk, x, false, tol_abs, a.cycles, L, &unconv,
                         true, own ? bk - mx0 : 0.0);
    } else {
        const float* Mf = reinterpret_cast<const float*>(a.M) + d.moff;
        // ||M - M32||_F <= 2^-24 ||M||_F (round to nearest, normal range) + N 2^-150 (subnormal range)
        const double dM = 0x1p-24 * sqrt(bsum1(dsq, L)) + n * 0x1p-150;
        bool done = false;
        for (; sweeps < a.max_refine; ++sweeps) {
            double mx = mx0;   // sweep 0: formed by the assembly
            if (sweeps > 0) {
                if (own) L.vs[t] = x;
                __syncthreads();
                mx = mv_free_sym(n, pc, L);   // every thread: it synchronises
            }
            const double r = own ? bk - mx : 0.0;
            const double rn = sqrt(bsum1(r * r, L));
            if (!(rn > tol_abs)) {
                done = true;
                break;
            }
            double dx = 0.0;
            iters += wg_gmres([&]() { return mv_tiles(Mf, n, L); }, n, r, dx, true,
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
    // One sweep over the unordered pairs: the cotangent sums of both rows of each pair (the image
    // sums too for the film).
    double* psy = reinterpret_cast<double*>(L.fa);   // 2 kBT doubles: sy | ix
    double* piy = reinterpret_cast<double*>(L.fb);
    {
        const int kr = sym_row(n);
        const NodeTrig tk = L.tr[kr];
        const double mk = L.vs[kr];
        if (pc.image) {
            double* const acc[4] = {L.part, psy, psy + kBT, piy};
            sym_sweep<4>(n, acc, [&](int k, int j, int, int, auto mirror, auto full, double (&row)[4], double (&col)[4]) {
                if (!decltype(full)::value && (k >= n || j >= n)) return;
                const NodeTrig tj = L.tr[j];
                const double mj = L.vs[j];
                const CotParts cp = cot_parts(tk, tj);
                const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
                const double w = mj * cp.id, wi = mj * cq.id;
                row[0] += w * cp.p;
                row[1] -= w * cp.q;
                row[2] += wi * cq.p;
                row[3] -= wi * cq.q;
                if constexpr (decltype(mirror)::value) {   // p, p' change sign with the pair's order, q changes, q' not
                    const double wk = mk * cp.id, wki = mk * cq.id;
                    col[0] -= wk * cp.p;
                    col[1] += wk * cp.q;
                    col[2] -= wki * cq.p;
                    col[3] -= wki * cq.q;
                }
            });
        } else {
            double* const acc[2] = {L.part, psy};
            sym_sweep<2>(n, acc, [&](int k, int j, int, int, auto mirror, auto full, double (&row)[2], double (&col)[2]) {
                if (!decltype(full)::value && (k >= n || j >= n)) return;
                const CotParts cp = cot_parts(tk, L.tr[j]);
                const double w = L.vs[j] * cp.id;
                row[0] += w * cp.p;
                row[1] -= w * cp.q;
                if constexpr (decltype(mirror)::value) {
                    const double wk = mk * cp.id;
                    col[0] -= wk * cp.p;
                    col[1] += wk * cp.q;
                }
            });
        }
    }
    double2 q1 = make_double2(0.0, 0.0), q2 = make_double2(0.0, 0.0);
    if (own) {
        const double sx = sym_total(L.part, n), sy = sym_total(psy, n);
        const double ix = pc.image ? sym_total(psy + kBT, n) : 0.0, iy = pc.image ? sym_total(piy, n) : 0.0;
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
