// sfh_kernels.hpp -- declarations of the gfx950 kernels of the Roberts stepper.
#pragma once
#include "sfh_common.hpp"

namespace sfh {

constexpr int kNB = 64;          // LU panel / tile width

// Stage control for the fused velocity + Bernoulli + RK4 kernel.
//   stage 0    : plain RHS evaluation, writes (q, dphi) into kZ[0], kP[0]
//   stage 1..3 : writes k_s into kZ[s-1], kP[s-1] and y_s = y0 + alpha k_s into (Zout, Pout)
//   stage 4    : y0 += alpha * (((k1 + 2 k2) + 2 k3) + k4) in place (alpha = dt / 6)
struct StageCtl {
    int stage;
    int pad;
    double alpha;
    const double2* Z0;
    const double* P0;
    double2* Zout;
    double* Pout;
    double2* kZ[3];
    double* kP[3];
    double2* qup;      // optional upper-fluid velocity output (may be null)
    const double* Pin; // phi of the stage's input state (read only when energy is set)
    double* energy;    // optional: 4 per node, the energy and flux terms of node g at 4 g
    // augmented optical drive: the delayed intensity D of the stage input, of y0 (updated in place
    // by stage 4), the stage output and the stage slopes (null without the augmented drive)
    const double* Din;
    double* D0;
    double* Dout;
    double* kD[3];
    int* nonfinite;    // optional: set to 1 when a stage slope of any node is NaN or infinite
    // timed drive: the stage's delayed intensity is dca * Din + dcb * I; dsave: store it in Dout
    double dca, dcb;
    int dsave;
    int pad_timed;
};

// Warm start of a solve by extrapolation over time steps: each RK stage keeps its solutions of the
// last kHistDepth steps (newest first, levels stride apart) and the next initial guess is the
// polynomial through the hv of them that exist, evaluated one step ahead:
//   hv = 1: u0;  2: 2 u0 - u1;  3: 3 (u0 - u1) + u2;  4: 4 (u0 + u2) - 6 u1 - u3;
//   5: 5 (u0 - u3) - 10 (u1 - u2) + u4   (error O(dt^hv); the coefficients are the binomial
//   coefficients of the hv-th difference, whose absolute sum 2^hv - 1 bounds how much the solve
//   tolerance left in the history is amplified: 31 at hv = 5).
constexpr int kHistDepth = 5;
__device__ inline double extrapolate_values(const double (&h)[kHistDepth], int hv) {
    if (hv >= 5) return 5.0 * (h[0] - h[3]) - 10.0 * (h[1] - h[2]) + h[4];
    if (hv == 4) return 4.0 * (h[0] + h[2]) - 6.0 * h[1] - h[3];
    if (hv == 3) return 3.0 * (h[0] - h[1]) + h[2];
    if (hv == 2) return 2.0 * h[0] - h[1];
    return h[0];
}
__device__ inline double extrapolate_history(const double* h, long long stride, int hv) {
    double v[kHistDepth];
#pragma unroll
    for (int l = 0; l < kHistDepth; ++l) v[l] = l < hv ? h[l * stride] : 0.0;
    return extrapolate_values(v, hv);
}
__device__ inline void push_history(double* h, long long stride, double x) {
#pragma unroll
    for (int l = kHistDepth - 1; l > 0; --l) h[l * stride] = h[(l - 1) * stride];
    h[0] = x;
}

struct GmArgs {
    const double* M;     // matrices (FP64, row-major per surface at SurfDesc::moff)
    const float* M32;    // the same in FP32 (mixed precision)
    const double* b;     // right-hand sides
    double* x;           // initial guess (written by zphi_circ_kernel, read only during the solve)
    double* sol;         // solution, written by the launch that finishes the solve
    double* V;           // Krylov basis, (m+1) * voff per surface
    double* w;           // two product vectors of total nodes each (even / odd steps)
    double* part;        // per-workgroup partial sums, part_stride per surface
    double* small;       // per surface: scales, coefficients, rotations, g, H
    int* ctl;            // per surface control block
    long long part_stride;
    long long small_stride;
    int m;               // maximal number of iterations (Krylov dimension)
    int total;           // nodes of all surfaces (stride between the two product vectors)
    double tol;
    // warm start by extrapolation over steps: stage s (1..4) keeps its solutions of the last
    // kHistDepth steps at hist + ((s - 1) kHistDepth + l) total, l = 0 newest; hvalid[s - 1]
    // counts them.  stage 0: plain warm start from x.
    double* hist;
    int* hvalid;
    int stage;
    int check_first;     // refinement solve: every launch checks for a finished solve before its matrix loads
};
constexpr int kPairRows = 4;        // default rows per workgroup of the N^2 pair kernels (assembly, velocity)
// ---- spectral.hip
// circulant derivative products: rows per workgroup (wave w owns rows w and w + 4), columns
// staged in LDS per round, and the kernel window one round needs
constexpr int kCircRowsPerWG = 8;
constexpr int kCircChunk = 2048;
constexpr int kCircWin = kCircChunk + kCircRowsPerWG - 1;
constexpr int kCircPer = kCircChunk / 256;               // staged samples per thread
constexpr int kCircWinPer = (kCircWin + 255) / 256;      // staged kernel entries per thread
// Window entry u holds c[(k0 - j0 - (kCircChunk - 1) + u) mod n]; row r, column c of the chunk
// read entry r - c + kCircChunk - 1.  circ_window gives the kernel indices of the entries
// u = threadIdx.x + 256 e a thread stages (entries beyond the window get valid indices too): one
// division, then steps of 256 mod n, each reduced by one conditional subtraction.
__device__ inline void circ_window(int k0, int j0, int n, int (&q)[kCircWinPer]) {
    int r = (k0 - j0 - (kCircChunk - 1) + (int)threadIdx.x) % n;
    if (r < 0) r += n;
    const int step = 256 % n;
#pragma unroll
    for (int e = 0; e < kCircWinPer; ++e) {
        q[e] = r;
        r += step;
        if (r >= n) r -= n;
    }
}
__global__ void zphi_circ_kernel(const SurfDesc* sd, const double* circ, const double2* Z, const double* phi,
                                 double2* Zp, double2* Zpp, double* b, double* work, Phys ph, int* status,
                                 NodeTrig* trig, GmArgs gm, int gm_on);
__global__ void mu_circ_kernel(const SurfDesc* sd, const double* circ, const double* mu, double2* mup);
// ---- deriv.hip
constexpr int kDerivThreads = 1024;
__global__ void plain_deriv_kernel(const SurfDesc* sd, const double2* tw_all, const double2* in,
                                   double2* out, int second, double scale, int lds_tw);

// ---- assemble.hip
// R = rows per workgroup: 1, 2 or 4 (instantiated in assemble.hip)
template <int R>
__global__ void assemble64_kernel(const SurfDesc* sd, const NodeTrig* trig, const double2* Z,
                                  const double2* Zp, const double2* Zpp, double* M, Phys ph,
                                  const int* only_if, GmArgs gm, int gm_on);
template <int R>
__global__ void assemble32_kernel(const SurfDesc* sd, const NodeTrig* trig, const double2* Z,
                                  const double2* Zp, const double2* Zpp, float* M, Phys ph, GmArgs gm);
template <int R>
__global__ void residual64_kernel(const SurfDesc* sd, const NodeTrig* trig, const double2* Z,
                                  const double2* Zp, const double2* Zpp, Phys ph, GmArgs gm);

// ---- velocity.hip
// R = rows per workgroup: 1, 2 or 4 (instantiated in velocity.hip)
template <int R>
__global__ void velocity_stage_kernel(const SurfDesc* sd, const NodeTrig* trig, const double2* Z,
                                      const double2* Zp, const double2* Zpp, const double* mu,
                                      const double2* mup, Phys ph, StageCtl ctl);
template <int R>
__global__ void velocity_gm1_kernel(const SurfDesc* sd, const NodeTrig* trig, const double2* Z,
                                    const double2* Zp, const double2* Zpp, double2* mup, Phys ph,
                                    StageCtl ctl, GmArgs gm, const double* circ);

// ---- lu.hip
struct LuArgs {
    double* A;        // n x n row-major, overwritten by L21 / U12 blocks (diagonal blocks kept)
    double* work;     // right-hand side, updated in place (b_i -= L21_i y_1)
    double* y;        // forward-substituted right-hand side L^{-1} P b
    double* x;        // solution
    double* uinv;     // per panel: U11^{-1} (64 x 64)
    int* flags;       // per panel ready flags for the back substitution
    int* status;      // [0] pivot check failed (needs interchanges), [1] spin timeout
    int n;
};
__global__ void lu_step_kernel(LuArgs a, int k0);
__global__ void lu_backsolve_kernel(LuArgs a);
__global__ void gepp_fallback_kernel(const SurfDesc* sd, double* M, const double* b, double* x,
                                     int* piv, int* status, int* counters);

// ---- gmres.hip
constexpr int kGmCtlInts = 32;   // ints per control block (holds the device struct)
template <typename T>
__global__ void gm_iter_kernel(GmArgs a, const SurfDesc* sd, int j);
__global__ void gm_close_mu_kernel(GmArgs a, const SurfDesc* sd, const double* circ, double2* mup, int i);

// ---- solver.hip: one trajectory frame [x(T), y(T), phi(T)] of the state
__global__ void record_frame_kernel(const double2* Z, const double* P, const double* D, double* dst, int total);
// state in the reference layout [Z(T), Phi(T) + 0i] <-> Z, P; and the right-hand side out
__global__ void state_split_kernel(const double2* state, double2* Z, double* P, int total);
__global__ void state_join_kernel(const double2* Z, const double* P, double2* state, int total);

// ---- batched.hip (persistent per-surface stage kernel, ensembles with every N <= kBatchMaxN)
constexpr int kBatchThreads = 1024;
constexpr int kBatchMaxN = 1024;
// The batched kernel stores a surface's matrix in 64 x 64 tiles (tile-skewed, sfh_sym.hpp): this
// many elements per surface.
constexpr int kTile = 64;
constexpr int kTileElems = kTile * kTile;
__host__ __device__ constexpr inline long long tiled_matrix_elems(int n) {
    const long long nb = (n + kTile - 1) / kTile;
    return nb * nb * kTileElems;
}
struct BatchArgs {
    const SurfDesc* sd;
    const int* order;          // workgroup -> surface, by decreasing N
    const double2* tw_all;
    const double2* Zin;        // stage input state
    const double* Pin;
    void* M;                   // per-surface column-major matrices (double, or float when mixed)
    double2* Zp;
    double2* Zpp;
    double2* mup;
    double* b;
    double* mu;                // warm start in, solution out
    int* stats;                // per surface: [iterations, unconverged solves, refinement sweeps, -]
    double tol;
    int mixed;
    int cycles;                // GMRES restart cycles (FP64)
    int max_refine;            // refinement sweeps (mixed)
    int nsurf;
    int total;                 // nodes of all surfaces
    int pad;
    // warm start by extrapolation over steps: stage s (1..4) keeps each surface's mu of the last
    // kHistDepth steps at hist + ((s - 1) kHistDepth + l) total, l = 0 newest;
    // hvalid[(s - 1) nsurf + i] counts them for surface i
    double* hist;
    int* hvalid;
};
__global__ void batched_stage_kernel(BatchArgs a, Phys ph, StageCtl ctl);

// ---- adjoint.hip (reverse mode of the water right-hand side, every N <= kBatchMaxN)
// Per-node intermediates kept in global scratch across the transpose solve.
struct AdjNode {
    double2 gzp, gzpp, gz, gw, p1, S;
    double2 pi1;   // film: sum over all j of mu_j (1 + cot((z_k - conj z_j)/2 + i h)^2)
    double mu, pad;
};
struct VjpArgs {
    const SurfDesc* sd;
    const double2* tw_all;
    const double2* Zin;        // state at which the Jacobian is taken
    const double* Pin;
    const double* mu_in;       // the forward mu at that state (null: solve for it)
    const double2* gA_Z;       // input cotangent = ca * gA + cb * gB (gB may be null)
    const double* gA_P;
    const double2* gB_Z;
    const double* gB_P;
    double ca, cb;
    double2* out_Z;            // J^T (input cotangent)
    double* out_P;
    double2* acc_Z;            // optional accumulator += out
    double* acc_P;
    AdjNode* scratch;
    int* stats;                // per surface: [iterations, unconverged, -, -]
    double tol;
    int cycles;
    int pad;
};
__global__ void rhs_vjp_kernel(VjpArgs a, Phys ph);

// ---- adjoint_mwg.hip (the reverse mode over many workgroups per surface)
struct AdjMArgs {
    const SurfDesc* sd;
    const NodeTrig* trig;      // half-angle table of the state
    const double2* Z;          // the state, its Z', Z'', mu and mu'
    const double2* Zp;
    const double2* Zpp;
    const double2* mup;
    const double* mu;
    const double2* gA_Z;       // input cotangent = ca * gA + cb * gB (gB may be null)
    const double* gA_P;
    const double2* gB_Z;
    const double* gB_P;
    double ca, cb;
    AdjNode* node;             // per-node intermediates
    double2* gwc;              // conj(gw)
    double2* gmup;             // cotangent of mu'
    double2* d1;               // D1 gmup (unscaled)
    double* gmu;               // cotangent of mu: right-hand side of M^T gb = gmu
    double* gb;                // its solution
    double2 *gam, *gamI, *gzp, *gzpp, *gz;
    double2* out_Z;
    double* out_P;
    double2* acc_Z;            // optional accumulator += out
    double* acc_P;
    double* quirk;             // water rho != 0 with the V1[1] quirk: rho gp, q2.x, q1.x per node (3 x total)
    int total;                 // nodes of all surfaces
};
constexpr int kAdjRows = 2;   // rows per workgroup of the reverse pair kernels
__global__ void adj_rho_quirk_kernel(AdjMArgs a);
__global__ void aug_vjp_kernel(const double2* Z, const double* gA_P, const double* gB_P, const double* gA_D,
                               const double* gB_D, double ca, double cb, double2* outZ, double2* accZ, double* outD,
                               double* accD, Phys ph, int total);
template <int R> __global__ void adj_pass_a_kernel(AdjMArgs a, Phys ph);
template <int R> __global__ void adj_pass_b_kernel(AdjMArgs a, Phys ph);
template <int R> __global__ void adj_pass_c_kernel(AdjMArgs a, Phys ph);
__global__ void adj_circ_d1_kernel(const SurfDesc* sd, const double* circ, const double2* v, double2* out);
__global__ void adj_gm_start_kernel(const SurfDesc* sd, const double* b, double* V, double* x, int* ctl, int m);
__global__ void adj_extrapolate_kernel(const double* ring, int head, int hv, double* x0, int total);
__global__ void adj_gm_start_warm_kernel(const SurfDesc* sd, const double* Mt, const double* b, const double* x0,
                                         double* V, double* x, int* ctl, int m);
__global__ void transpose_kernel(const SurfDesc* sd, const double* M, double* Mt);
__global__ void adj_gamma_kernel(AdjMArgs a, Phys ph);
__global__ void adj_final_kernel(AdjMArgs a, const double* circ, const double2* sd1gb);

}  // namespace sfh
