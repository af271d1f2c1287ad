// sfh_common.hpp -- shared device/host definitions of the gfx950 Roberts boundary-integral stepper.
//
// Layout in HBM (one "ensemble" = S independent periodic surfaces of N_s nodes each):
//   vectors  : surfaces concatenated, node i of surface s at voff[s] + i
//              Z, Z', Z'', mu', q  : double2 (re, im)      16 B/node
//              phi, b (= Re Phi'), mu, dphi/dt : double    8 B/node
//   matrices : surface s at moff[s], row-major N_s x N_s (row k = collocation node, column j =
//              source node), i.e. element (k, j) of the reference's system (createM.cuh:52).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfh {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Per-surface descriptor, uploaded once per ensemble.
struct SurfDesc {
    int n;          // nodes
    int log2m;      // n = r * 2^log2m with r in {1, 3}
    int r3;         // 1 if n = 3 * 2^log2m
    int voff;       // vector offset (nodes)
    long long moff; // matrix offset (elements)
    int twoff;      // offset of this n's twiddle table (complex entries, n of them)
    int coff;       // offset of this n's circulant derivative kernels (3 n doubles: Re c1, Im c1, c2)
};

// Physics selector passed by value to kernels.
//   kind 0: water, createMKernel / createVelocityMatrices / compute_rhs_phi_expression
//   kind 1: helium film, createFiniteDepthMKernel / createHeliumVelocityMatrices /
//           compute_rhs_helium_phi_expression (+ kappa and expansion variants)
struct Phys {
    int kind;
    int image;          // helium with finite depth: image terms on
    int use_exp;
    int exp_order;
    int rho_quirk;      // water Bernoulli reads V1[1] (createM.cuh:104)
    int need_upper;     // water with rho != 0 needs the upper-fluid velocity
    double rho, kappa, depth, U, g;
    // optical drive (HeliumDrivenAutonomousProblem, AugmentedBoundaryIntegrator), nondimensional:
    // dPhi/dt += damping v + strength I(y, x) [+ D], dD/dt = beta I - D / tau when augmented;
    // timed: the delayed intensity of HeliumWithOptomechanicalDrivingProblem, an exponential
    // integrator over the stage times (dPhi/dt += damping v + beta D + strength I)
    int drive;
    int augmented;
    int timed;
    int pad_drive;
    double damping, strength, detuning, gamma, G, max_intensity, x0, sigma, beta, tau;
};

// ------------------------------------------------------------------ complex helpers
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 ctimes_i(double2 a) { return make_double2(-a.y, a.x); }
__device__ inline double2 cdiv(double2 a, double2 b) {
    double d = 1.0 / (b.x * b.x + b.y * b.y);
    return make_double2((a.x * b.x + a.y * b.y) * d, (a.y * b.x - a.x * b.y) * d);
}
__device__ inline double2 crecip(double2 b) {
    double d = 1.0 / (b.x * b.x + b.y * b.y);
    return make_double2(b.x * d, -b.y * d);
}

// cot(a + i b) = (sin a cos a - i sinh b cosh b) / (sin^2 a + sinh^2 b)  (from DLMF 4.21.40 with
// cosh 2b - cos 2a = 2 (sin^2 a + sinh^2 b), which removes the cancellation near the poles).
// sinh/cosh come from one expm1 so that small |b| keeps full relative accuracy.
__device__ inline double2 cot64(double a, double b) {
    double sa, ca;
    sincos(a, &sa, &ca);
    double ab = fabs(b);
    if (ab < 20.0) {
        double E = expm1(ab);
        double ep = E + 1.0;
        double iep = 1.0 / ep;
        double sh = 0.5 * E * (E + 2.0) * iep;   // sinh |b|
        double ch = 0.5 * (ep + iep);            // cosh b
        double id = 1.0 / (sa * sa + sh * sh);
        return make_double2(sa * ca * id, -copysign(sh * ch, b) * id);
    }
    double t = exp(-2.0 * ab);
    return make_double2(4.0 * sa * ca * t, -copysign(1.0 + 2.0 * t, b));
}

__device__ inline float2 cot32(float a, float b) {
    float sa, ca;
    sincosf(a, &sa, &ca);
    float ab = fabsf(b);
    if (ab < 10.0f) {
        float E = expm1f(ab);
        float ep = E + 1.0f;
        float iep = 1.0f / ep;
        float sh = 0.5f * E * (E + 2.0f) * iep;
        float ch = 0.5f * (ep + iep);
        float id = 1.0f / (sa * sa + sh * sh);
        return make_float2(sa * ca * id, -copysignf(sh * ch, b) * id);
    }
    float t = expf(-2.0f * ab);
    return make_float2(4.0f * sa * ca * t, -copysignf(1.0f + 2.0f * t, b));
}

// 1/d for the pair kernels: the hardware reciprocal (v_rcp_f64) refined by two Newton-Raphson
// steps r <- r + r (1 - d r), each of which squares the relative error, so the result is within a
// few ulp of 1/d (not correctly rounded).  5 instructions against 12 for an IEEE division.  Only
// for the positive, normal-range denominators sin^2 + sinh^2 of distinct nodes.
__device__ inline double fast_rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    return r;
}

// Per-node table for pairwise cotangents: sin(x/2), cos(x/2), exp(y/2) / 2, exp(-y/2).
// With it cot((z_k - z_j)/2) = (sa ca - i sh ch) / (sa^2 + sh^2) needs no transcendental call:
//   sa = sin((x_k - x_j)/2) = s_k c_j - c_k s_j,  ca = c_k c_j + s_k s_j,
//   b = (y_k - y_j)/2:  e^b / 2 = e_k ei_j,  e^-b / 2 = e_j ei_k  (the factor 1/2 is in e, exactly),
//   sh = sinh b = e_k ei_j - e_j ei_k,  ch = cosh b = e_k ei_j + e_j ei_k.
// The subtractions lose relative accuracy only in proportion to eps / |z_k - z_j|, the same order
// as the rounding already present in forming x_k - x_j and y_k - y_j.
struct NodeTrig {
    double s, c, e, ei;
};

__device__ inline NodeTrig node_trig(double2 z) {
    NodeTrig t;
    sincos(0.5 * z.x, &t.s, &t.c);
    const double E = exp(0.5 * z.y);
    t.e = 0.5 * E;
    t.ei = 1.0 / E;
    return t;
}

// A pair cotangent in parts: cot = (p id, -q id) with p = sin a cos a, q = sinh b cosh b and
// id = 1 / (sin^2 a + sinh^2 b), so a consumer folds its own factors into one product with id.
struct CotParts {
    double p, q, id;
};

// cot((z_k - z_j) / 2)
__device__ inline CotParts cot_parts(const NodeTrig& k, const NodeTrig& j) {
    const double sa = k.s * j.c - k.c * j.s;
    const double ca = k.c * j.c + k.s * j.s;
    const double ep = k.e * j.ei, em = k.ei * j.e;
    const double sh = ep - em, ch = ep + em;
    return {sa * ca, sh * ch, fast_rcp(sa * sa + sh * sh)};
}
__device__ inline double2 cot_pair(const NodeTrig& k, const NodeTrig& j) {
    const CotParts c = cot_parts(k, j);
    return make_double2(c.p * c.id, -c.q * c.id);
}

// cot((z_k - conj z_j) / 2 + i h), eh = exp(h), ehi = exp(-h): b = (y_k + y_j)/2 + h > 0,
// e^b / 2 = e_k e_j (2 eh), e^-b / 2 = ei_k ei_j (ehi / 2).
__device__ inline CotParts cot_image_parts(const NodeTrig& k, const NodeTrig& j, double eh, double ehi) {
    const double sa = k.s * j.c - k.c * j.s;
    const double ca = k.c * j.c + k.s * j.s;
    const double ep = k.e * j.e * (2.0 * eh), em = k.ei * j.ei * (0.5 * ehi);
    const double sh = ep - em, ch = ep + em;
    return {sa * ca, sh * ch, fast_rcp(sa * sa + sh * sh)};
}
__device__ inline double2 cot_image(const NodeTrig& k, const NodeTrig& j, double eh, double ehi) {
    const CotParts c = cot_image_parts(k, j, eh, ehi);
    return make_double2(c.p * c.id, -c.q * c.id);
}

// ------------------------------------------------------------------ wave reductions
// v taken from another lane of the same row of 16 by a DPP control (both halves of the double).
template <int CTL>
__device__ inline double dpp_double(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double read_lane(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Sum over the 64 lanes of a wave (every lane active), the same value in every lane: inside each
// row of 16 by DPP (quad_perm xor 1, xor 2, half-row mirror, row mirror; a lane adds the
// partner's partial sum to its own, and both orders of the two operands give the same double),
// then the four row sums read as scalars and added in a fixed order.
__device__ inline double wave_sum(double v) {
    v += dpp_double<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_double<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_double<0x141>(v);   // row_half_mirror
    v += dpp_double<0x140>(v);   // row_mirror
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Sums of NV per-thread values over a 256-thread workgroup, into out[0..NV) (LDS; valid after the
// call).  NV >= 4: the values go through LDS transposed, P = 256 / NV threads add NV entries each
// and a P-lane butterfly finishes, so no value takes a full 64-lane shuffle chain; NV < 4: one
// wave sum per value and the four waves' results added.  Fixed order: the result does not depend
// on timing.  red holds max(NV, 4) * 256 doubles.
template <int NV>
__device__ inline void rows_reduce(const double (&v)[NV], double* red, double* out) {
    static_assert(NV >= 1 && NV <= 256 && (NV & (NV - 1)) == 0, "NV: a power of two in [1, 256]");
    const int t = threadIdx.x;
    if constexpr (NV < 4) {
        const int lane = t & 63, w = t >> 6;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const double s = wave_sum(v[q]);
            if (lane == 0) red[q * 4 + w] = s;
        }
        __syncthreads();
        if (t < NV) out[t] = red[t * 4] + red[t * 4 + 1] + red[t * 4 + 2] + red[t * 4 + 3];
        __syncthreads();
    } else {
        constexpr int P = 256 / NV;
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q * 256 + t] = v[q];
        __syncthreads();
        const int q = t / P, p = t % P;
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < NV; ++e) s += red[q * 256 + p + P * e];
        for (int o = P / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (p == 0) out[q] = s;
        __syncthreads();
    }
}

}  // namespace sfh
