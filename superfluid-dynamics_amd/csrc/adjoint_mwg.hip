// adjoint_mwg.hip -- the reverse mode of adjoint.hip spread over many workgroups per surface, for
// the multi-workgroup solver paths (single large surfaces, uniform batches).  The reverse pass is the
// same (its derivation heads adjoint.hip); it is split where a pass needs every node of the previous
// one:
//   adj_pass_a_kernel   per row: S, P1 (the film's SI, PI1), then the node-local adjoint of the
//                       velocity and Bernoulli terms (gw, the cotangent of mu', the local gmu, gZ',
//                       gZ'' and the Bernoulli term's gz)
//   adj_circ_d1_kernel  D1 of the cotangent of mu' (D1^H = -D1)
//   adj_pass_b_kernel   per row: P3 (the film's Q) -> gmu, the right-hand side of M^T gb = gmu
//   transpose_kernel    M^T of the re-assembled M; the GMRES launch sequence of the forward path then
//                       solves for gb (its closing launch also forms s D1 gb)
//   adj_gamma_kernel    the matrix terms and the pair weights gamma (the film's gammaI)
//   adj_pass_c_kernel   per row: P2 (the film's PI2) -> gz
//   adj_final_kernel    gZ += -s D1 gZ' + s^2 D2 gZ'', gPhi = -s Re D1 gb; output, accumulation
// The pair kernels take R rows per workgroup over all columns, as the forward velocity kernel does.
#include "sfh_gmres.hpp"
#include "sfh_physics.hpp"

namespace sfh {

namespace {

struct PairK {
    double c, ci, eh, ehi;
    int image;
};

__device__ inline PairK pair_k(const Phys& ph) {
    PairK p;
    p.c = ph.kind == 0 ? 0.25 * (1 - ph.rho) / kPi : 0.25 / kPi;
    p.ci = 0.25 / kPi;
    p.eh = exp(ph.depth);
    p.ehi = exp(-ph.depth);
    p.image = ph.image;
    return p;
}

}  // namespace

template <int R>
__global__ __launch_bounds__(256) void adj_pass_a_kernel(AdjMArgs a, Phys ph) {
    constexpr int NV = 8 * R;   // per row: S, P1, SI, PI1 (complex)
    __shared__ double red[NV * 256];
    __shared__ double sums[NV];
    const SurfDesc d = a.sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int k0 = blockIdx.x * R;
    if (k0 >= n) return;
    const int t = threadIdx.x;
    const PairK pk = pair_k(ph);
    NodeTrig tk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tk[r] = a.trig[off + min(k0 + r, n - 1)];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int j = t; j < n; j += 256) {
        const NodeTrig tj = a.trig[off + j];
        const double mj = a.mu[off + j];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = k0 + r;
            if (k >= n) break;
            double* v = acc + 8 * r;
            if (pk.image) {
                const double2 e = cot_image(tk[r], tj, pk.eh, pk.ehi);
                v[4] += mj * e.x;
                v[5] += mj * e.y;
                v[6] += mj * (1.0 + e.x * e.x - e.y * e.y);
                v[7] += mj * (2.0 * e.x * e.y);
            }
            if (j == k) continue;
            const double2 c = cot_pair(tk[r], tj);
            v[0] += mj * c.x;
            v[1] += mj * c.y;
            v[2] += mj * (1.0 + c.x * c.x - c.y * c.y);
            v[3] += mj * (2.0 * c.x * c.y);
        }
    }
    rows_reduce<NV>(acc, red, sums);
    if (t >= R || k0 + t >= n) return;
    const int g = off + k0 + t;
    const double* sv = sums + 8 * t;
    const double q4 = 0.25 / kPi;
    const double2 zp = a.Zp[g], zpp = a.Zpp[g], mup = a.mup[g];
    const double mu = a.mu[g];
    const double2 izp = crecip(zp), izp2 = cmul(izp, izp);
    const double2 V2 = ctimes_i(cscale(izp, 1.0 / (2.0 * kPi)));
    const double2 V1 = cadd(ctimes_i(cscale(cmul(zpp, izp2), -q4)), cscale(izp, 0.5));
    double2 w = cadd(cadd(cmul(V2, mup), cscale(V1, mu)), ctimes_i(make_double2(-q4 * sv[0], -q4 * sv[1])));
    if (pk.image) w = cadd(w, ctimes_i(make_double2(q4 * sv[4], q4 * sv[5])));
    const double2 q1 = cconj(w), q2 = cconj(csub(w, cscale(izp, mu)));   // lower and upper velocities
    double2 kqz = cscale(a.gA_Z[g], a.ca);
    double kp = a.ca * a.gA_P[g];
    if (a.gB_Z) {
        kqz = cadd(kqz, cscale(a.gB_Z[g], a.cb));
        kp += a.cb * a.gB_P[g];
    }
    // under the V1[1] quirk node 1's x velocity is not known here: its terms are added by
    // adj_rho_quirk_kernel from the values kept in a.quirk
    const bool quirk = ph.kind == 0 && ph.rho != 0.0 && ph.rho_quirk && n > 1;
    double2 gq1, gq2;
    bernoulli_vjp_q(ph, q1, q2, 0.0, quirk, kqz, kp, gq1, gq2);
    if (quirk) {
        a.quirk[g] = ph.rho * kp;
        a.quirk[a.total + g] = q2.x;
        a.quirk[2 * a.total + g] = q1.x;
    }
    const VelVjp vj = velocity_vjp(zp, zpp, mup, mu, gq1, gq2);
    AdjNode A;
    A.gzp = vj.gzp;
    A.gzpp = vj.gzpp;
    curvature_vjp(ph, zp, zpp, kp, A.gzp, A.gzpp);
    A.gz = bernoulli_gz(ph, a.Z[g].x, a.Z[g].y, kp);
    A.gw = vj.gw;
    A.p1 = make_double2(sv[2], sv[3]);
    A.S = make_double2(sv[0], sv[1]);
    A.pi1 = make_double2(sv[6], sv[7]);
    A.mu = mu;
    A.pad = 0.0;
    a.node[g] = A;
    a.gwc[g] = cconj(vj.gw);
    a.gmup[g] = vj.gmup;
    a.gmu[g] = vj.gmu;
}

// The V1[1] quirk's terms after pass A (water, rho != 0): every node's Bernoulli term reads node 1's
// lower x velocity q1x_1, so node k's gq2.x gains -rho gp_k q1x_1 and node 1's gq1.x gains
// -rho sum_k gp_k q2.x_k.  Both enter the node-local cotangents linearly (velocity_vjp) and are added
// to them before any later pass reads them.  Grid (ceil(N / 256), surfaces); workgroup 0 of a surface
// also forms the sum and applies node 1's term after its own.
__global__ __launch_bounds__(256) void adj_rho_quirk_kernel(AdjMArgs a) {
    __shared__ double red[4];
    const int s = blockIdx.y;
    const SurfDesc d = a.sd[s];
    const int n = d.n, t = threadIdx.x, k = blockIdx.x * 256 + t;
    if (n <= 1 || blockIdx.x * 256 >= n) return;
    const double q1x1 = a.quirk[2 * a.total + d.voff + 1];
    auto add = [&](int g, double2 gq1, double2 gq2) {
        const VelVjp vj = velocity_vjp(a.Zp[g], a.Zpp[g], a.mup[g], a.mu[g], gq1, gq2);
        AdjNode& A = a.node[g];
        A.gw = cadd(A.gw, vj.gw);
        A.gzp = cadd(A.gzp, vj.gzp);
        A.gzpp = cadd(A.gzpp, vj.gzpp);
        a.gwc[g] = cadd(a.gwc[g], cconj(vj.gw));
        a.gmup[g] = cadd(a.gmup[g], vj.gmup);
        a.gmu[g] += vj.gmu;
    };
    if (k < n) add(d.voff + k, make_double2(0.0, 0.0), make_double2(-a.quirk[d.voff + k] * q1x1, 0.0));
    if (blockIdx.x != 0) return;
    double v = 0.0;
    for (int j = t; j < n; j += 256) v -= a.quirk[d.voff + j] * a.quirk[a.total + d.voff + j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 1) add(d.voff + 1, make_double2((red[0] + red[1]) + (red[2] + red[3]), 0.0), make_double2(0.0, 0.0));
}

// The augmented drive's delayed intensity D (AugmentedBoundaryIntegrator.cuh:32-36, createM.cuh:151-169):
// dPhi/dt gains + D and dD/dt = beta I(x, y) - D / tau, both linear in D and local to the node.  With
// the stage cotangents kp (of dPhi/dt) and kD (of dD/dt), each ca * gA + cb * gB:
//   gD = kp - kD / tau,  gz += kD beta (dI/dx + i dI/dy),
// added to the reverse pass of the two-field right-hand side (outZ, and accZ when accumulating).
__global__ __launch_bounds__(256) void aug_vjp_kernel(const double2* __restrict__ Z, const double* gA_P,
                                                      const double* gB_P, const double* gA_D, const double* gB_D,
                                                      double ca, double cb, double2* outZ, double2* accZ,
                                                      double* outD, double* accD, Phys ph, int total) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= total) return;
    double kp = ca * gA_P[g], kD = ca * gA_D[g];
    if (gB_P) kp += cb * gB_P[g];
    if (gB_D) kD += cb * gB_D[g];
    const double oD = kp - kD / ph.tau;
    outD[g] = oD;
    if (accD) accD[g] += oD;
    const double2 gi = intensity_grad(ph, Z[g].x, Z[g].y);
    const double2 gz = make_double2(kD * ph.beta * gi.x, kD * ph.beta * gi.y);
    outZ[g] = cadd(outZ[g], gz);
    if (accZ) accZ[g] = cadd(accZ[g], gz);
}

// out_k = sum_j c1[(k - j) mod n] v_j (the first-derivative circulant, unscaled), one row per wave.
__global__ __launch_bounds__(256) void adj_circ_d1_kernel(const SurfDesc* __restrict__ sd,
                                                          const double* __restrict__ circ,
                                                          const double2* __restrict__ v,
                                                          double2* __restrict__ out) {
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= n) return;   // the whole wave
    const double* c1r = circ + d.coff;
    const double* c1i = c1r + n;
    double ax = 0.0, ay = 0.0;
    for (int j = lane; j < n; j += 64) {
        int q = k - j;
        if (q < 0) q += n;
        const double cr = c1r[q], ci = c1i[q];
        const double2 x = v[off + j];
        ax += cr * x.x - ci * x.y;
        ay += cr * x.y + ci * x.x;
    }
    ax = wave_sum(ax);
    ay = wave_sum(ay);
    if (lane == 0) out[off + k] = make_double2(ax, ay);
}

template <int R>
__global__ __launch_bounds__(256) void adj_pass_b_kernel(AdjMArgs a, Phys ph) {
    constexpr int NV = 2 * R;   // per row: Im P3, Im Q
    __shared__ double red[(NV < 4 ? 4 : NV) * 256];
    __shared__ double sums[NV];
    const SurfDesc d = a.sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int j0 = blockIdx.x * R;
    if (j0 >= n) return;
    const int t = threadIdx.x;
    const PairK pk = pair_k(ph);
    NodeTrig tj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = a.trig[off + min(j0 + r, n - 1)];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int k = t; k < n; k += 256) {
        const NodeTrig tk = a.trig[off + k];
        const double2 wk = a.gwc[off + k];   // conj(gw_k)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = j0 + r;
            if (j >= n) break;
            if (pk.image) {
                const double2 e = cot_image(tj[r], tk, pk.eh, pk.ehi);
                acc[2 * r + 1] += wk.x * e.y - wk.y * e.x;   // Im(gw_k e_jk)
            }
            if (k == j) continue;
            const double2 c = cot_pair(tj[r], tk);
            acc[2 * r] += wk.x * c.y + wk.y * c.x;           // Im(conj(gw_k) c_jk)
        }
    }
    rows_reduce<NV>(acc, red, sums);
    if (t >= R || j0 + t >= n) return;
    const int g = off + j0 + t;
    const double q4 = 0.25 / kPi, s = 2.0 * kPi / n;
    a.gmu[g] += -s * a.d1[g].x - q4 * sums[2 * t] - q4 * sums[2 * t + 1];
}

// GMRES start for M^T gb = gmu from a zero guess: V[0] = r0 = gmu, x = 0, the control block re-armed.
__global__ __launch_bounds__(256) void adj_gm_start_kernel(const SurfDesc* __restrict__ sd,
                                                           const double* __restrict__ b, double* __restrict__ V,
                                                           double* __restrict__ x, int* __restrict__ ctl, int m) {
    const SurfDesc d = sd[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < d.n) {
        V[(long long)(m + 1) * d.voff + i] = b[d.voff + i];
        x[d.voff + i] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        GmCtl* c = reinterpret_cast<GmCtl*>(ctl) + blockIdx.y;
        c->done = 0;
        c->refine_iters = 0;
        c->iters = 0;
    }
}

// The warm start of a transposed solve: x0 = the extrapolation (quartic once five are kept) of the
// solutions of the same stage over the previous reverse steps, kept in a ring of kHistDepth vectors
// (slot (head - l) mod kHistDepth holds the solution l steps back).
__global__ __launch_bounds__(256) void adj_extrapolate_kernel(const double* __restrict__ ring, int head, int hv,
                                                              double* __restrict__ x0, int total) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    double h[kHistDepth];
#pragma unroll
    for (int l = 0; l < kHistDepth; ++l) {
        const int slot = (head - l + kHistDepth) % kHistDepth;
        h[l] = l < hv ? ring[(long long)slot * total + i] : 0.0;
    }
    x0[i] = extrapolate_values(h, hv);
}

// The GMRES launch sequence on M^T from the warm start x0: V[0] = b - M^T x0 (one wave per row of the
// explicit M^T, 4 rows per workgroup), x = x0, and the control block re-armed.
__global__ __launch_bounds__(256) void adj_gm_start_warm_kernel(const SurfDesc* __restrict__ sd,
                                                                const double* __restrict__ Mt,
                                                                const double* __restrict__ b,
                                                                const double* __restrict__ x0,
                                                                double* __restrict__ V, double* __restrict__ x,
                                                                int* __restrict__ ctl, int m) {
    const SurfDesc d = sd[blockIdx.y];
    const int n = d.n, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        GmCtl* c = reinterpret_cast<GmCtl*>(ctl) + blockIdx.y;
        c->done = 0;
        c->refine_iters = 0;
        c->iters = 0;
    }
    if (k >= n) return;   // the whole wave
    const double* row = Mt + d.moff + (long long)k * n;
    const double* xs = x0 + d.voff;
    double acc = 0.0;
    for (int j = lane; j < n; j += 64) acc += row[j] * xs[j];
    acc = wave_sum(acc);
    if (lane == 0) {
        V[(long long)(m + 1) * d.voff + k] = b[d.voff + k] - acc;
        x[d.voff + k] = xs[k];
    }
}

// Mt = M^T per surface, through 32 x 32 LDS tiles.
__global__ __launch_bounds__(256) void transpose_kernel(const SurfDesc* __restrict__ sd, const double* __restrict__ M,
                                                        double* __restrict__ Mt) {
    __shared__ double tile[32][33];
    const SurfDesc d = sd[blockIdx.z];
    const int n = d.n;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    if (bx >= n || by >= n) return;   // the whole workgroup
    const double* A = M + d.moff;
    double* B = Mt + d.moff;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int row = by + r, col = bx + tx;
        if (row < n && col < n) tile[r][tx] = A[(long long)row * n + col];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int row = bx + r, col = by + tx;
        if (row < n && col < n) B[(long long)row * n + col] = tile[tx][r];
    }
}

// The matrix terms (off-diagonal gZ'_k -= i c gb_k conj(S_k), the diagonal M_kk's Z' and Z'') and the
// pair weights gamma_k = (i/8pi) conj(gw_k) - (i c/2) gb_k Z'_k, gz_k += conj(gamma_k P1_k); the film's
// gammaI_k = (i/8pi) gw_k - (i ci/2) gb_k, gz_k += gammaI_k conj(PI1_k).
__global__ __launch_bounds__(256) void adj_gamma_kernel(AdjMArgs a, Phys ph) {
    const SurfDesc d = a.sd[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= d.n) return;
    const int g = d.voff + i;
    const PairK pk = pair_k(ph);
    const double q4 = 0.25 / kPi;
    const AdjNode A = a.node[g];
    const double gb = a.gb[g];
    const double2 zp = a.Zp[g], zpp = a.Zpp[g];
    double2 gzp = cadd(A.gzp, ctimes_i(cscale(cconj(A.S), -pk.c * gb)));
    const double akk = -gb * A.mu;
    const double2 gzpp = cadd(A.gzpp, ctimes_i(cscale(crecip(cconj(zp)), pk.c * akk)));
    gzp = cadd(gzp, ctimes_i(cscale(cconj(cdiv(zpp, cmul(zp, zp))), -pk.c * akk)));
    const double2 gam = ctimes_i(csub(cscale(cconj(A.gw), 0.5 * q4), cscale(zp, 0.5 * pk.c * gb)));
    double2 gz = cadd(A.gz, cconj(cmul(gam, A.p1)));
    if (pk.image) {
        const double2 gI = ctimes_i(csub(cscale(A.gw, 0.5 * q4), make_double2(0.5 * pk.ci * gb, 0.0)));
        gz = cadd(gz, cmul(gI, cconj(A.pi1)));
        a.gamI[g] = gI;
    }
    a.gam[g] = gam;
    a.gzp[g] = gzp;
    a.gzpp[g] = gzpp;
    a.gz[g] = gz;
}

template <int R>
__global__ __launch_bounds__(256) void adj_pass_c_kernel(AdjMArgs a, Phys ph) {
    constexpr int NV = 2 * R;   // per row: P2 (complex)
    __shared__ double red[(NV < 4 ? 4 : NV) * 256];
    __shared__ double sums[NV];
    const SurfDesc d = a.sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int j0 = blockIdx.x * R;
    if (j0 >= n) return;
    const int t = threadIdx.x;
    const PairK pk = pair_k(ph);
    NodeTrig tj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = a.trig[off + min(j0 + r, n - 1)];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int k = t; k < n; k += 256) {
        const NodeTrig tk = a.trig[off + k];
        const double2 gk = a.gam[off + k];
        const double2 gi = pk.image ? a.gamI[off + k] : make_double2(0.0, 0.0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = j0 + r;
            if (j >= n) break;
            if (pk.image) {
                const double2 e = cot_image(tj[r], tk, pk.eh, pk.ehi);
                const double ex = 1.0 + e.x * e.x - e.y * e.y, ey = 2.0 * e.x * e.y;
                acc[2 * r] += gi.x * ex - gi.y * ey;
                acc[2 * r + 1] += gi.x * ey + gi.y * ex;
            }
            if (k == j) continue;
            const double2 c = cot_pair(tj[r], tk);
            const double ex = 1.0 + c.x * c.x - c.y * c.y, ey = 2.0 * c.x * c.y;
            acc[2 * r] += gk.x * ex - gk.y * ey;
            acc[2 * r + 1] += gk.x * ey + gk.y * ex;
        }
    }
    rows_reduce<NV>(acc, red, sums);
    if (t >= R || j0 + t >= n) return;
    const int g = off + j0 + t;
    const double mj = a.mu[g];
    a.gz[g] = csub(a.gz[g], cscale(make_double2(sums[2 * t], -sums[2 * t + 1]), mj));
}

// gZ_k = gz_k - s (D1 gZ')_k + s^2 (D2 gZ'')_k, gPhi_k = -Re(s D1 gb)_k (the closing GMRES launch left
// s D1 gb in sd1gb); one row per wave.
__global__ __launch_bounds__(256) void adj_final_kernel(AdjMArgs a, const double* __restrict__ circ,
                                                        const double2* __restrict__ sd1gb) {
    const SurfDesc d = a.sd[blockIdx.y];
    const int n = d.n, off = d.voff;
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= n) return;   // the whole wave
    const double* c1r = circ + d.coff;
    const double* c1i = c1r + n;
    const double* c2 = c1i + n;
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (int j = lane; j < n; j += 64) {
        int q = k - j;
        if (q < 0) q += n;
        const double cr = c1r[q], ci = c1i[q], cc = c2[q];
        const double2 u = a.gzp[off + j], v = a.gzpp[off + j];
        ax += cr * u.x - ci * u.y;
        ay += cr * u.y + ci * u.x;
        bx += cc * v.x;
        by += cc * v.y;
    }
    ax = wave_sum(ax);
    ay = wave_sum(ay);
    bx = wave_sum(bx);
    by = wave_sum(by);
    if (lane != 0) return;
    const int g = off + k;
    const double s = 2.0 * kPi / n, s2 = s * s;
    const double2 gz = a.gz[g];
    const double2 gZ = make_double2(gz.x - s * ax + s2 * bx, gz.y - s * ay + s2 * by);
    const double gP = -sd1gb[g].x;
    a.out_Z[g] = gZ;
    a.out_P[g] = gP;
    if (a.acc_Z) {
        a.acc_Z[g] = cadd(a.acc_Z[g], gZ);
        a.acc_P[g] += gP;
    }
}

template __global__ void adj_pass_a_kernel<kAdjRows>(AdjMArgs, Phys);
template __global__ void adj_pass_b_kernel<kAdjRows>(AdjMArgs, Phys);
template __global__ void adj_pass_c_kernel<kAdjRows>(AdjMArgs, Phys);

}  // namespace sfh
