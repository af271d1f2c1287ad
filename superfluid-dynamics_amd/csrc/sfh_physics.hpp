// sfh_physics.hpp -- per-node physics shared by the fused velocity kernel and the persistent
// per-surface stage kernel: velocity from the cot sums, Bernoulli right-hand side, RK4 stage update.
#pragma once
#include "sfh_kernels.hpp"

namespace sfh {

__device__ inline double bernoulli(const Phys& ph, double y, double2 q1, double2 q2, double2 q1_quirk,
                                   double2 zp, double2 zpp) {
    if (ph.kind == 0) {
        const double a1 = q1.x * q1.x + q1.y * q1.y;
        const double a2 = q2.x * q2.x + q2.y * q2.y;
        const double dot = q1_quirk.x * q2.x + q1.y * q2.y;    // createM.cuh:104 reads V1[1].real()
        return -(1 + ph.rho) * ph.g * y + 0.5 * a1 + 0.5 * ph.rho * a2 - ph.rho * dot;
    }
    const double h = ph.depth;
    if (ph.use_exp) {
        const double kin = 0.5 * q1.x * q1.x + 0.5 * q1.y * q1.y;
        double v = 0.0;
        switch (ph.exp_order) {
        case 3: v += -10.0 / 3.0 * (y * y * y) / (h * h); [[fallthrough]];
        case 2: v += 2.0 * (y * y) / h; [[fallthrough]];
        case 1: v += -y; [[fallthrough]];
        default: break;
        }
        return ph.g * v + kin;
    }
    if (ph.kappa != 0.0) {
        const double a1 = q1.x * q1.x + q1.y * q1.y;
        const double s = zp.x * zp.x + zp.y * zp.y;
        const double curv = (zp.x * zpp.y - zp.y * zpp.x) / (s * sqrt(s));
        const double u = 1.0 + y / h;
        return ph.g * (20.447761896665416 * h / 3.0 * (1.0 / (u * u * u) - 1)) + 0.5 * a1 + ph.kappa * curv;
    }
    const double vdw = ph.g * (h / 3.0);
    const double u = 1.0 + y / h;
    return vdw / (u * u * u) - vdw + 0.5 * q1.x * q1.x + 0.5 * q1.y * q1.y;
}

// d(dPhi/dt)/dy of the Bernoulli right-hand side at height y, for the adjoint: water -(1 + rho) g,
// the helium film's van der Waals term -g (1 + y/h)^-4 (-20.447761896665416 g (1 + y/h)^-4 in the
// branch with surface tension) or the derivative of its expansion.
__device__ inline double bernoulli_dy(const Phys& ph, double y) {
    if (ph.kind == 0) return -(1 + ph.rho) * ph.g;
    const double h = ph.depth;
    if (ph.use_exp) {
        double v = 0.0;
        switch (ph.exp_order) {
        case 3: v += -10.0 * (y * y) / (h * h); [[fallthrough]];
        case 2: v += 4.0 * y / h; [[fallthrough]];
        case 1: v += -1.0; [[fallthrough]];
        default: break;
        }
        return ph.g * v;
    }
    const double u = 1.0 + y / h;
    const double a = ph.kappa != 0.0 ? 20.447761896665416 : 1.0;
    return -a * ph.g / ((u * u) * (u * u));
}

// The surface-tension term kappa curv, curv = (x' y'' - y' x'') / |Z'|^3, of the helium Bernoulli
// right-hand side: its cotangents w.r.t. Z' and Z'' (g = d/dRe + i d/dIm) for a cotangent kp of
// dPhi/dt, added to gzp and gzpp (nothing without the term).
__device__ inline void curvature_vjp(const Phys& ph, double2 zp, double2 zpp, double kp, double2& gzp, double2& gzpp) {
    if (ph.kind == 0 || ph.use_exp || ph.kappa == 0.0) return;
    const double s = zp.x * zp.x + zp.y * zp.y;
    const double r3 = 1.0 / (s * sqrt(s)), r5 = r3 / s;
    const double cr = zp.x * zpp.y - zp.y * zpp.x;
    const double w = ph.kappa * kp;
    gzpp = cadd(gzpp, make_double2(-zp.y * r3 * w, zp.x * r3 * w));
    gzp = cadd(gzp, make_double2((zpp.y * r3 - 3.0 * cr * zp.x * r5) * w, (-zpp.x * r3 - 3.0 * cr * zp.y * r5) * w));
}

// Velocities at node k from the cot sums over j != k
//   S = sum_j mu_j cot((z_k - z_j)/2) = sx + i sy,  Si = sum_j mu_j cot((z_k - conj z_j)/2 + i h) = ix + i iy
//   w = V2_k mu'_k + V1_kk mu_k + i (-1/4pi) S [+ i (1/4pi) Si]
//   V2_k = i / (2 pi Z'_k),  V1_kk = i (-1/4pi) Z''_k / Z'_k^2 [+ i (1/4pi) cot(i (y_k + h))] +- 1/(2 Z'_k)
// q1 (lower fluid, + sign) and q2 (upper fluid, - sign) are conj(w) (WaterVelocities.cuh:38-107).
__device__ inline void velocity_from_sums(const Phys& ph, double2 zp, double2 zpp, double2 mup, double mk,
                                          double yk, double sx, double sy, double ix, double iy,
                                          double2& q1, double2& q2) {
    const double q4 = 1.0 / (4.0 * kPi);
    double2 w = ctimes_i(make_double2(-q4 * sx, -q4 * sy));
    if (ph.image) w = cadd(w, ctimes_i(make_double2(q4 * ix, q4 * iy)));
    double2 v1 = ctimes_i(cscale(cdiv(zpp, cmul(zp, zp)), -q4));
    if (ph.image) v1 = cadd(v1, ctimes_i(cscale(cot64(0.0, yk + ph.depth), q4)));
    const double2 half_inv = cscale(crecip(zp), 0.5);
    const double2 v2 = ctimes_i(crecip(cscale(zp, 2.0 * kPi)));
    const double2 base = cadd(cadd(cmul(v2, mup), w), cscale(v1, mk));
    q1 = cconj(cadd(base, cscale(half_inv, mk)));
    q2 = cconj(csub(base, cscale(half_inv, mk)));
}

// Reverse mode of the node-local part of velocity_from_sums (the cot sums' terms go to the pair
// passes).  With w = V2 mu' + V1 mu + (sums), V1 = -(i/4pi) Z''/Z'^2, and the jump mu/(2 Z') added for
// the lower velocity q1 = conj(w + mu/(2 Z')) and subtracted for the upper one q2 = conj(w - mu/(2 Z')),
// the cotangents (gq1, gq2) of (q1, q2) give
//   gw = conj(gq1 + gq2) (the cotangent of w and of every sum), gh = conj(gq1 - gq2) (of the jump),
//   gmu' = conj(V2) gw,  gmu = Re(conj(V1) gw) + Re(conj(1/(2Z')) gh),
//   gZ' = conj(dV2/dZ') conj(mu') gw + conj(dV1/dZ') mu gw + conj(-1/(2Z'^2)) mu gh,
//   gZ'' = conj(dV1/dZ'') mu gw.
// Linear in (gq1, gq2), with no constant term.
struct VelVjp {
    double2 gw, gmup, gzp, gzpp;
    double gmu;
};
__device__ inline VelVjp velocity_vjp(double2 zp, double2 zpp, double2 mup, double mu, double2 gq1, double2 gq2) {
    const double q4 = 0.25 / kPi;
    const double2 izp = crecip(zp), izp2 = cmul(izp, izp);
    const double2 V2 = ctimes_i(cscale(izp, 1.0 / (2.0 * kPi)));
    const double2 V1 = ctimes_i(cscale(cmul(zpp, izp2), -q4));
    const double2 dV2 = ctimes_i(cscale(izp2, -1.0 / (2.0 * kPi)));
    const double2 dV1p = ctimes_i(cscale(cmul(zpp, cmul(izp2, izp)), 1.0 / (2.0 * kPi)));
    const double2 dV1pp = ctimes_i(cscale(izp2, -q4));
    VelVjp r;
    r.gw = cconj(cadd(gq1, gq2));
    const double2 gh = cconj(csub(gq1, gq2));
    r.gmup = cmul(cconj(V2), r.gw);
    r.gmu = cmul(cconj(V1), r.gw).x + cmul(cconj(cscale(izp, 0.5)), gh).x;
    r.gzp = cadd(cadd(cmul(cconj(dV2), cmul(cconj(mup), r.gw)), cmul(cconj(dV1p), cscale(r.gw, mu))),
                 cmul(cconj(cscale(izp2, -0.5)), cscale(gh, mu)));
    r.gzpp = cmul(cconj(dV1pp), cscale(r.gw, mu));
    return r;
}

// Cotangents (gq1, gq2) of the lower and upper velocities from those of the right-hand side
// (kqz for dz/dt = q1, kp for dPhi/dt), with the autonomous drive's damping term damping q1.y
// (createM.cuh:138-149).  Water (createM.cuh:96-107):
//   dPhi/dt = -(1 + rho) g y + |q1|^2/2 + rho |q2|^2/2 - rho (q1q.x q2.x + q1.y q2.y),
// q1q = q1 but, under the V1[1] quirk, node 1's lower velocity (createM.cuh:104).  The quirk's
// q1q.x term belongs to node 1: the caller passes q1q.x of node 1 (or 0 and adds -rho kp q1q.x to
// gq2.x later) and adds -rho kp q2.x of every node to node 1's gq1.x.  The film has no upper term.
__device__ inline void bernoulli_vjp_q(const Phys& ph, double2 q1, double2 q2, double q1qx, bool quirk,
                                       double2 kqz, double kp, double2& gq1, double2& gq2) {
    gq1 = cadd(kqz, cscale(q1, kp));
    if (ph.drive) gq1.y += kp * ph.damping;   // the drive's damping term, damping q1.y
    gq2 = make_double2(0.0, 0.0);
    if (ph.kind == 0 && ph.rho != 0.0) {
        const double r = ph.rho * kp;
        gq1 = csub(gq1, make_double2(quirk ? 0.0 : r * q2.x, r * q2.y));
        gq2 = make_double2(r * (q2.x - (quirk ? q1qx : q1.x)), r * (q2.y - q1.y));
    }
}

// LightIntensity::compute_intensity (LightIntensity.cuh:16-29)
__device__ inline double intensity(const Phys& ph, double y, double x) {
    const double df = ph.detuning - ph.G * y;
    const double dx = x - ph.x0;
    const double hg = ph.gamma / 2.0;
    return 0.25 * (ph.gamma * ph.gamma) * ph.max_intensity / (df * df + hg * hg) *
           exp(-(dx * dx) / (2.0 * (ph.sigma * ph.sigma)));
}

// Optical drive on the Bernoulli right-hand side of node g, in the reference's order:
// add_optical_field_drive_terms_no_time_depence (createM.cuh:138-149: damping v, then strength I),
// then add_delayed_intensity_phi_rhs (createM.cuh:151-159: + D) and the delayed intensity's own
// slope calculate_intensity_delayed_rhs (createM.cuh:161-169: beta I - D / tau).
__device__ inline double drive_terms(const Phys& ph, const StageCtl& ctl, int g, double x, double y, double2 q1,
                                     double dphi, double& dD) {
    const double inten = intensity(ph, y, x);
    dD = 0.0;
    if (ph.timed) {
        // add_optical_field_drive_terms (createM.cuh:119-136) with DelayedIntensityTerm's update
        // (DelayedIntensityTerm.cuh:20-37): D = a Dsaved + beta tau (1 - a) I, the coefficients of
        // the stage's time formed on the host; the first stage of a step saves D
        const double D = ctl.dca * ctl.Din[g] + ctl.dcb * inten;
        if (ctl.dsave) ctl.Dout[g] = D;
        dphi += ph.damping * q1.y;
        dphi += ph.beta * D;
        dphi += ph.strength * inten;
        return dphi;
    }
    dphi += ph.damping * q1.y;
    dphi += ph.strength * inten;
    if (ph.augmented) {
        const double D = ctl.Din[g];
        dphi += D;
        dD = ph.beta * inten - 1.0 / ph.tau * D;
    }
    return dphi;
}

// The cotangent of the node position z = x + i y from dPhi/dt's explicit dependence on it (kp its
// cotangent): i (d dPhi/dt / dy) kp from the restoring term, plus for the autonomous drive
// strength kp (dI/dx + i dI/dy), I = (gamma^2/4) I_max / ((delta - G y)^2 + gamma^2/4) exp(-(x - x0)^2 / (2 sigma^2)):
// dI/dy = 2 G (delta - G y) I / ((delta - G y)^2 + gamma^2/4), dI/dx = -(x - x0) I / sigma^2.
__device__ inline double2 intensity_grad(const Phys& ph, double x, double y) {
    const double inten = intensity(ph, y, x);
    const double df = ph.detuning - ph.G * y, hg = ph.gamma / 2.0;
    return make_double2(inten * (-(x - ph.x0) / (ph.sigma * ph.sigma)), inten * (2.0 * ph.G * df / (df * df + hg * hg)));
}
__device__ inline double2 bernoulli_gz(const Phys& ph, double x, double y, double kp) {
    double2 g = make_double2(0.0, bernoulli_dy(ph, y) * kp);
    if (ph.drive) {
        const double2 gi = intensity_grad(ph, x, y);
        g.x += ph.strength * kp * gi.x;
        g.y += ph.strength * kp * gi.y;
    }
    return g;
}

// RK4 stage bookkeeping of node g (AutonomousRungeKuttaStepper.cuh:338-375, utilities.cuh:78-83),
// in two halves so that a kernel can load a node's stage inputs long before it has the node's
// velocity: stage_load reads what the stage combines, stage_store writes the stage's results.
struct StageIn {
    double2 z0, k1, k2, k3;
    double p0, p1, p2, p3;
};
__device__ inline StageIn stage_load(const StageCtl& ctl, int g) {
    StageIn s{};
    if (ctl.stage >= 1) {
        s.z0 = ctl.Z0[g];
        s.p0 = ctl.P0[g];
    }
    if (ctl.stage >= 4) {
        s.k1 = ctl.kZ[0][g];
        s.k2 = ctl.kZ[1][g];
        s.k3 = ctl.kZ[2][g];
        s.p1 = ctl.kP[0][g];
        s.p2 = ctl.kP[1][g];
        s.p3 = ctl.kP[2][g];
    }
    return s;
}
// The delayed intensity D of the augmented drive takes the same stage combination as phi.
__device__ inline void stage_store_delayed(const StageCtl& ctl, int g, double dD) {
    if (ctl.stage == 0) {
        ctl.kD[0][g] = dD;
    } else if (ctl.stage <= 3) {
        ctl.kD[ctl.stage - 1][g] = dD;
        ctl.Dout[g] = ctl.D0[g] + ctl.alpha * dD;
    } else {
        const double K = ctl.kD[0][g] + 2.0 * ctl.kD[1][g] + 2.0 * ctl.kD[2][g] + dD;
        ctl.D0[g] = ctl.D0[g] + ctl.alpha * K;
    }
}
__device__ inline void stage_store(const StageCtl& ctl, int g, double2 q1, double dphi, const StageIn& s) {
    // every writer stores the same value, so the unordered plain store is enough
    if (ctl.nonfinite && !(isfinite(q1.x) && isfinite(q1.y) && isfinite(dphi))) *ctl.nonfinite = 1;
    if (ctl.stage == 0) {
        ctl.kZ[0][g] = q1;
        ctl.kP[0][g] = dphi;
    } else if (ctl.stage <= 3) {
        ctl.kZ[ctl.stage - 1][g] = q1;
        ctl.kP[ctl.stage - 1][g] = dphi;
        ctl.Zout[g] = make_double2(s.z0.x + ctl.alpha * q1.x, s.z0.y + ctl.alpha * q1.y);
        ctl.Pout[g] = s.p0 + ctl.alpha * dphi;
    } else {
        const double Kx = s.k1.x + 2.0 * s.k2.x + 2.0 * s.k3.x + q1.x;
        const double Ky = s.k1.y + 2.0 * s.k2.y + 2.0 * s.k3.y + q1.y;
        const double Kp = s.p1 + 2.0 * s.p2 + 2.0 * s.p3 + dphi;
        const_cast<double2*>(ctl.Z0)[g] = make_double2(s.z0.x + ctl.alpha * Kx, s.z0.y + ctl.alpha * Ky);
        const_cast<double*>(ctl.P0)[g] = s.p0 + ctl.alpha * Kp;
    }
}
// Per-node terms of EnergyContainer::CalculateEnergy and VolumeFlux (Energies.cuh:136-204),
// summed per surface and scaled on the host: kinetic (phi + U(1+rho)x/2)(-Y' u + X' v)
// - (U/2)((u + rho u) X' + (v + rho v) Y' + U(1-rho) X'/2) y (KineticEnergy passes the lower
// velocity for the upper one), potential y^2 X' (water) or (1 + y/h)^-2 - 1 (helium), |Z'|, and
// v X' + u Y'.  q = u + i v is the lower velocity, Z' = X' + i Y'.
__device__ inline void energy_store(const Phys& ph, const StageCtl& ctl, int g, double x, double y, double phi,
                                    double2 zp, double2 q) {
    const double U = ph.U, rho = ph.rho;
    const double kin = (phi + 0.5 * U * (1.0 + rho) * x) * (-zp.y * q.x + zp.x * q.y) -
                       0.5 * U * ((q.x + rho * q.x) * zp.x + (q.y + rho * q.y) * zp.y + 0.5 * U * (1.0 - rho) * zp.x) * y;
    double pot;
    if (ph.kind == 0) {
        pot = y * y * zp.x;
    } else {
        const double r = 1.0 + y / ph.depth;
        pot = 1.0 / (r * r) - 1.0;
    }
    double2* e = reinterpret_cast<double2*>(ctl.energy) + 2 * (long long)g;
    e[0] = make_double2(kin, pot);
    e[1] = make_double2(sqrt(zp.x * zp.x + zp.y * zp.y), q.y * zp.x + q.x * zp.y);
}

__device__ inline void stage_update(const StageCtl& ctl, int g, double2 q1, double dphi) {
    stage_store(ctl, g, q1, dphi, stage_load(ctl, g));
}

}  // namespace sfh
