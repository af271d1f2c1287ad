// adjoint.hip -- hand-written reverse mode (vector-Jacobian product) of the Roberts right-hand
// side for water and the helium film, one workgroup per surface (every N <= 1024).
//
// Forward map (BaseBoundaryIntegralCalculator::runTimeStep, BaseBoundaryIntegrator.cuh:137-314):
//   Z' = s D1 (Z - L) + s,  Z'' = s^2 D2 (Z - L),  b = s Re D1 Phi      (s = 2 pi / N)
//   M_kj = c Im(Z'_k c_kj) (j != k),  M_kk = (1 + rho)/2 + c Im(Z''_k / Z'_k),  c_kj = cot((z_k - z_j)/2),
//   c = (1 - rho)/4pi
//   mu = M^{-1} b,  mu' = s D1 mu
//   w_k = V2_k mu'_k + V1_k mu_k - (i/4pi) S_k,  S_k = sum_{j!=k} mu_j c_kj,  q = conj(w)
//   V2 = i / (2 pi Z'),  V1 = -(i/4pi) Z'' / Z'^2 + 1 / (2 Z')  (q2, the upper velocity: - 1 / (2 Z'))
//   dPhi/dt = -(1 + rho) g y + |q|^2 / 2 + rho |q2|^2 / 2 - rho (q1q.x q2.x + q.y q2.y), q1q = q, or
//   node 1's q under the V1[1] quirk
// (createM.cuh:43-63,96-107; WaterVelocities.cuh:38-70,205-242).  D1, D2 are the spectral
// derivative operators of the reference's multiplier table; every D1 multiplier is imaginary and
// every D2 multiplier real, so D1^H = -D1 and D2^H = D2.
//
// Cotangents are complex numbers g with dL = Re sum conj(g) dz.  Given (gq, gp) for (q, dPhi/dt):
//   the Bernoulli term gives the cotangents gq1, gq2 of q and q2 (bernoulli_vjp_q; under the quirk
//   node 1's x part collects -rho gp q2.x of every node), gw = conj(gq1 + gq2);  the node-local
//   parts of w and of the jump +-mu/(2Z') give gZ', gZ'', gmu (velocity_vjp), and gz = i (d dPhi/dt/dy) gp;
//   gmu += s Re D1^H (conj V2 gw)  + Re((i/4pi) P3),  P3_j = sum_k conj(gw_k) c_jk;
//   M^T gb = gmu (GMRES, matrix-free transposed product);
//   gZ' += -i c gb conj(S) (off-diagonal M), the diagonal of M adds to gZ', gZ'';
//   pair terms (d c_kj = -(1 + c_kj^2)(dz_k - dz_j)/2 in both the velocity sum and M):
//     gamma_k = (i/8pi) conj(gw_k) - (i c/2) gb_k Z'_k
//     gz_k += conj(gamma_k P1_k),  P1_k = sum_j mu_j (1 + c_kj^2)
//     gz_j -= mu_j conj(P2_j),     P2_j = sum_k gamma_k (1 + c_jk^2)
//   gPhi = s Re D1^H gb,  gZ += s D1^H gZ' + s^2 D2^H gZ''.
// Helium film with finite depth h (createFiniteDepthMKernel createM.cuh:65-92, the helium velocities
// WaterVelocities.cuh:72-107, compute_rhs_helium_phi_expression createM.cuh:109-117): with the image
// cotangent e_kj = cot((z_k - conj z_j)/2 + i h) over ALL j (j = k gives the diagonal terms cot(i(y_k + h)))
//   M_kj += -ci Im e_kj,  w_k += (i/4pi) sum_j mu_j e_kj,  dPhi/dt = g h/3 (u^-3 - 1) + |q|^2/2, u = 1 + y/h
// (ci = 1/4pi; the film's M image term has no Z'_k, as in the reference).  e_kj = -conj(e_jk), so the
// image part of M is symmetric, and with d e_kj = -(1 + e_kj^2)(dz_k - conj dz_j)/2:
//   gmu_j += -Im(Q_j)/4pi,  Q_j = sum_k gw_k e_jk
//   gammaI_k = (i/8pi) gw_k - (i ci/2) gb_k
//   gz_k += gammaI_k conj(PI1_k),  PI1_k = sum_j mu_j (1 + e_kj^2)
//   gz_j -= mu_j conj(PI2_j),      PI2_j = sum_k gammaI_k (1 + e_jk^2)
//   gz += i (d/dy dPhi/dt) gp:  -g u^-4 (van der Waals) or the expansion's derivative; the
//   surface-tension term kappa (x'y'' - y'x'')/|Z'|^3 adds its Z' and Z'' derivatives to gZ', gZ''.
// The RK4 step adjoint around it (stage recomputation, cotangent combination) is host code.
#include "sfh_wg.hpp"

namespace sfh {

// (M^T v)_j = M_jj v_j - c Im(sum_{k != j} v_k Z'_k c_jk), v in L.vs.
__device__ inline double mv_free_t(int n, int G, const PairConst& pc, BLds& L) {
    const int t = threadIdx.x;
    double val = 0.0;
    if (t < G * n) {
        const int j = t % n, g = t / n;
        const NodeTrig tj = L.tr[j];
        double uy = 0.0, ui = 0.0;   // Im sum v_k Z'_k c_jk, and the symmetric image part Im sum v_k e_jk
        for (int k = g; k < n; k += G) {
            if (k == j) continue;
            const double vk = L.vs[k];
            const double2 zk = L.zp[k];
            const double2 ct = cot_pair(tj, L.tr[k]);
            uy += vk * (zk.x * ct.y + zk.y * ct.x);
            if (pc.image) ui += vk * cot_image(tj, L.tr[k], pc.eh, pc.ehi).y;
        }
        val = -pc.c * uy - pc.ci * ui;
    }
    double w = gsum(val, n, G, L);
    if (t < n) w += L.diag[t] * L.vs[t];
    return w;
}

__global__ __launch_bounds__(1024) void rhs_vjp_kernel(VjpArgs a, Phys ph) {
    __shared__ BLds L;
    const int s = blockIdx.x;
    const SurfDesc d = a.sd[s];
    const int n = d.n, off = d.voff, t = threadIdx.x;
    const bool own = t < n;
    const int G = kBT / n > 0 ? kBT / n : 1;
    const int act = t < G * n;
    const int kr = act ? t % n : 0, grp = act ? t / n : 0;
    const double sc = 2.0 * kPi / n, sc2 = sc * sc;
    const double q4 = 0.25 / kPi;
    AdjNode* an = a.scratch + off;

    // ------------------------------------------------ forward recomputation
    const NodeVals nv = wg_derivatives(d, a.tw_all, a.Zin, a.Pin, ph, L);
    const PairConst pc = wg_diagonal(ph, nv, n, L);
    double mu = 0.0;
    int its = 0, unconv = 0;
    if (a.mu_in) {
        mu = own ? a.mu_in[off + t] : 0.0;
    } else {
        const double bn = sqrt(bsum1(own ? nv.b * nv.b : 0.0, L));
        its += wg_gmres([&]() { return mv_free(n, G, pc, L); }, n, nv.b, mu, true, a.tol * bn, a.cycles, L,
                        &unconv);
    }
    if (own) L.fa[lds_index(d, t)] = make_double2(mu, 0.0);
    wg_first_derivative(d, L, false);
    double2 mup = make_double2(0.0, 0.0);
    if (own) {
        const double2 v = L.fa[lds_index(d, t)];
        mup = make_double2(v.x * sc, v.y * sc);
        L.vs[t] = mu;
    }
    __syncthreads();
    // pass A: S_k = sum mu_j c_kj,  P1_k = sum mu_j (1 + c_kj^2) (j != k); film: SI_k = sum mu_j e_kj,
    // PI1_k = sum mu_j (1 + e_kj^2) over all j
    double sx = 0.0, sy = 0.0, p1x = 0.0, p1y = 0.0, six = 0.0, siy = 0.0, q1x = 0.0, q1y = 0.0;
    if (act) {
        const NodeTrig tk = L.tr[kr];
        for (int j = grp; j < n; j += G) {
            const double mj = L.vs[j];
            if (pc.image) {
                const double2 e = cot_image(tk, L.tr[j], pc.eh, pc.ehi);
                six += mj * e.x;
                siy += mj * e.y;
                q1x += mj * (1.0 + e.x * e.x - e.y * e.y);
                q1y += mj * (2.0 * e.x * e.y);
            }
            if (j == kr) continue;
            const double2 c = cot_pair(tk, L.tr[j]);
            sx += mj * c.x;
            sy += mj * c.y;
            p1x += mj * (1.0 + c.x * c.x - c.y * c.y);
            p1y += mj * (2.0 * c.x * c.y);
        }
    }
    sx = gsum(sx, n, G, L);
    sy = gsum(sy, n, G, L);
    p1x = gsum(p1x, n, G, L);
    p1y = gsum(p1y, n, G, L);
    if (pc.image) {
        six = gsum(six, n, G, L);
        siy = gsum(siy, n, G, L);
        q1x = gsum(q1x, n, G, L);
        q1y = gsum(q1y, n, G, L);
    }

    // ------------------------------------------------ node-local adjoint of velocity + Bernoulli
    const bool quirk = ph.kind == 0 && ph.rho != 0.0 && ph.rho_quirk && n > 1;
    double2 q1 = make_double2(0.0, 0.0), q2 = q1, kqz = q1;
    double kp = 0.0;
    if (own) {
        const double2 zp = nv.zp, zpp = nv.zpp;
        const double2 izp = crecip(zp), izp2 = cmul(izp, izp);
        const double2 V2 = ctimes_i(cscale(izp, 1.0 / (2.0 * kPi)));
        const double2 V1 = cadd(ctimes_i(cscale(cmul(zpp, izp2), -q4)), cscale(izp, 0.5));
        double2 w = cadd(cadd(cmul(V2, mup), cscale(V1, mu)), ctimes_i(make_double2(-q4 * sx, -q4 * sy)));
        if (pc.image) w = cadd(w, ctimes_i(make_double2(q4 * six, q4 * siy)));
        q1 = cconj(w);                         // lower velocity
        q2 = cconj(csub(w, cscale(izp, mu)));  // upper velocity (the jump mu/(2Z') subtracted)
        kqz = cscale(a.gA_Z[off + t], a.ca);
        kp = a.ca * a.gA_P[off + t];
        if (a.gB_Z) {
            kqz = cadd(kqz, cscale(a.gB_Z[off + t], a.cb));
            kp += a.cb * a.gB_P[off + t];
        }
    }
    double q1qx = 0.0;   // node 1's lower x velocity (the V1[1] quirk)
    if (quirk) {
        __syncthreads();
        if (t == 1) L.scal[2] = q1.x;
        __syncthreads();
        q1qx = L.scal[2];
    }
    double2 gq1, gq2;
    bernoulli_vjp_q(ph, q1, q2, q1qx, quirk, kqz, kp, gq1, gq2);
    if (quirk) {   // node 1's x velocity enters every node's Bernoulli term
        const double d = bsum1(own ? -ph.rho * kp * q2.x : 0.0, L);
        if (t == 1) gq1.x += d;
    }
    double2 gw = make_double2(0.0, 0.0);
    double gmu = 0.0;
    if (own) {
        const VelVjp vj = velocity_vjp(nv.zp, nv.zpp, mup, mu, gq1, gq2);
        gw = vj.gw;
        double2 gzp = vj.gzp, gzpp = vj.gzpp;
        curvature_vjp(ph, nv.zp, nv.zpp, kp, gzp, gzpp);
        gmu = vj.gmu;
        AdjNode& A = an[t];
        A.gzp = gzp;
        A.gzpp = gzpp;
        A.gz = bernoulli_gz(ph, nv.z.x, nv.z.y, kp);   // dPhi/dt's own dependence on z
        A.pi1 = make_double2(q1x, q1y);
        A.gw = gw;
        A.p1 = make_double2(p1x, p1y);
        A.mu = mu;
        A.S = make_double2(sx, sy);
        L.fa[lds_index(d, t)] = vj.gmup;
        L.fb[t] = cconj(gw);
    }
    // gmu += s Re(D1^H gmu')
    wg_first_derivative(d, L, true);
    if (own) gmu += sc * L.fa[lds_index(d, t)].x;
    // pass B: P3_j = sum_k conj(gw_k) c_jk ;  gmu_j += Re((i/4pi) P3_j); film: Q_j = sum_k gw_k e_jk
    // over all k, gmu_j += -Im(Q_j) / 4pi
    double p3x = 0.0, p3y = 0.0, qy = 0.0;
    if (act) {
        const NodeTrig tj = L.tr[kr];
        for (int k = grp; k < n; k += G) {
            const double2 wk = L.fb[k];   // conj(gw_k)
            if (pc.image) {
                const double2 e = cot_image(tj, L.tr[k], pc.eh, pc.ehi);
                qy += wk.x * e.y - wk.y * e.x;   // Im(gw_k e_jk), gw_k = conj(wk)
            }
            if (k == kr) continue;
            const double2 c = cot_pair(tj, L.tr[k]);
            p3x += wk.x * c.x - wk.y * c.y;
            p3y += wk.x * c.y + wk.y * c.x;
        }
    }
    p3x = gsum(p3x, n, G, L);
    p3y = gsum(p3y, n, G, L);
    if (pc.image) qy = gsum(qy, n, G, L);
    if (own) gmu += -q4 * p3y - q4 * qy;

    // ------------------------------------------------ transpose solve M^T gb = gmu
    double gb = 0.0;
    {
        const double gn = sqrt(bsum1(own ? gmu * gmu : 0.0, L));
        its += wg_gmres([&]() { return mv_free_t(n, G, pc, L); }, n, gmu, gb, true, a.tol * gn, a.cycles, L,
                        &unconv);
    }

    // ------------------------------------------------ matrix terms and pair weights gamma
    double2 gz = make_double2(0.0, 0.0), gzp = make_double2(0.0, 0.0), gzpp = make_double2(0.0, 0.0);
    double muk = 0.0;
    if (own) {
        const AdjNode A = an[t];
        muk = A.mu;
        const double2 zp = nv.zp, zpp = nv.zpp;
        const double2 S = A.S;
        gzp = A.gzp;
        gzpp = A.gzpp;
        gz = A.gz;
        // off-diagonal M:  gZ'_k += -i c gb_k conj(S_k)
        gzp = cadd(gzp, ctimes_i(cscale(cconj(S), -pc.c * gb)));
        // diagonal M_kk = 1/2 + c Im(Z''/Z'), weight a_kk = -gb_k mu_k
        const double akk = -gb * muk;
        gzpp = cadd(gzpp, ctimes_i(cscale(crecip(cconj(zp)), pc.c * akk)));
        gzp = cadd(gzp, ctimes_i(cscale(cconj(cdiv(zpp, cmul(zp, zp))), -pc.c * akk)));
        // gamma_k = (i/8pi) conj(gw_k) - (i c / 2) gb_k Z'_k
        const double2 gam = ctimes_i(csub(cscale(cconj(A.gw), 0.5 * q4), cscale(zp, 0.5 * pc.c * gb)));
        // gz_k += conj(gamma_k P1_k)
        gz = cadd(gz, cconj(cmul(gam, A.p1)));
        L.fb[t] = gam;
        L.vs[t] = muk;
        if (pc.image) {   // gammaI_k = (i/8pi) gw_k - (i ci/2) gb_k;  gz_k += gammaI_k conj(PI1_k)
            const double2 gI = ctimes_i(csub(cscale(A.gw, 0.5 * q4), make_double2(0.5 * pc.ci * gb, 0.0)));
            gz = cadd(gz, cmul(gI, cconj(A.pi1)));
            L.fa[t] = gI;
        }
    }
    __syncthreads();
    // pass C: P2_j = sum_k gamma_k (1 + c_jk^2);  gz_j -= mu_j conj(P2_j); film: PI2_j = sum_k
    // gammaI_k (1 + e_jk^2) over all k, gz_j -= mu_j conj(PI2_j)
    double p2x = 0.0, p2y = 0.0;
    if (act) {
        const NodeTrig tj = L.tr[kr];
        for (int k = grp; k < n; k += G) {
            if (pc.image) {
                const double2 gk = L.fa[k];
                const double2 e = cot_image(tj, L.tr[k], pc.eh, pc.ehi);
                const double ex = 1.0 + e.x * e.x - e.y * e.y, ey = 2.0 * e.x * e.y;
                p2x += gk.x * ex - gk.y * ey;
                p2y += gk.x * ey + gk.y * ex;
            }
            if (k == kr) continue;
            const double2 gk = L.fb[k];
            const double2 c = cot_pair(tj, L.tr[k]);
            const double ex = 1.0 + c.x * c.x - c.y * c.y, ey = 2.0 * c.x * c.y;
            p2x += gk.x * ex - gk.y * ey;
            p2y += gk.x * ey + gk.y * ex;
        }
    }
    p2x = gsum(p2x, n, G, L);
    p2y = gsum(p2y, n, G, L);
    if (own) gz = csub(gz, cscale(make_double2(p2x, -p2y), muk));

    // ------------------------------------------------ through the derivative operators
    // gPhi = s Re(D1^H gb)
    __syncthreads();
    if (own) L.fa[lds_index(d, t)] = make_double2(gb, 0.0);
    wg_first_derivative(d, L, true);
    double gphi = 0.0;
    if (own) gphi = sc * L.fa[lds_index(d, t)].x;
    // gZ += s D1^H gZ' + s^2 D2^H gZ''  (two forward transforms, one inverse)
    __syncthreads();
    if (own) {
        L.fa[lds_index(d, t)] = gzp;
        L.fb[lds_index(d, t)] = gzpp;
    }
    __syncthreads();
    fft_dit(L.fa, d, -1, L.tw);
    fft_dit(L.fb, d, -1, L.tw);
    __syncthreads();
    for (int i = t; i < n; i += kBT) {
        const double2 c1 = first_mult(L.fa[i], i, n);
        const double2 c2 = second_mult(L.fb[i], i, n);
        L.fa[i] = make_double2(-sc * c1.x + sc2 * c2.x, -sc * c1.y + sc2 * c2.y);
    }
    __syncthreads();
    fft_dif(L.fa, d, +1, L.tw);
    __syncthreads();
    if (own) {
        gz = cadd(gz, L.fa[lds_index(d, t)]);
        a.out_Z[off + t] = gz;
        a.out_P[off + t] = gphi;
        if (a.acc_Z) {
            a.acc_Z[off + t] = cadd(a.acc_Z[off + t], gz);
            a.acc_P[off + t] += gphi;
        }
    }
    if (t == 0) {
        a.stats[4 * s] = its;
        a.stats[4 * s + 1] += unconv;
    }
}

}  // namespace sfh
