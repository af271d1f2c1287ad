// lu.hip -- dense FP64 solve M mu = b for the dipole strengths (one large surface per launch
// grid, uniform batches on gridDim.z).
//
// Replaces MatrixSolver::solve (CuSuperHelium/MatrixSolver.cuh:113-172): cusolverDnDgetrf +
// cusolverDnDgetrs (LU with partial pivoting), or getrfBatched/getrsBatched with a host sync.
//
// Algorithm.  Right-looking blocked LU with 64-wide panels, one launch (lu_step_kernel) per panel
// p (columns k0..k0+kb): every panel workgroup applies panel p-1's update to the diagonal block
// and to its own block, factors the kb x kb diagonal block A11 = L11 U11 in LDS and inverts both
// factors; then workgroup 0 stores U11^{-1} and y_p = L11^{-1} b_p, each "row" workgroup forms
// L21_i = A21_i U11^{-1} and b_i -= L21_i y_p, each "column" workgroup forms
// U12_j = L11^{-1} A12_j; the other workgroups of the launch apply panel p-1's update
// A22 -= L21 U12 (v_mfma_f64_16x16x4 on 64 x 64 tiles) to the rest of the trailing matrix.
// lu_backsolve_kernel then runs x_p = U_pp^{-1} (y_p - sum_{q>p} U_pq x_q), one workgroup per
// block row, chained by per-block ready flags.
//
// Row interchanges.  Partial pivoting (getrf: first row of maximal |a| in the current column)
// leaves the rows in place exactly when every multiplier satisfies |l_ij| <= 1.  Both the
// diagonal-block elimination and the L21 blocks check that condition; if it ever fails,
// status[0] is set and gepp_fallback_kernel re-solves the freshly re-assembled system with
// explicit interchanges.  The result is therefore the GEPP solution in every case, with the
// interchange-free path taken for the diagonally dominated second-kind systems of this method.
#include "sfh_kernels.hpp"

namespace sfh {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int LD = kNB + 1;   // LDS leading dimension of 64 x 64 FP64 blocks

__device__ inline v4d mfma_f64(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// acc[i][j] (+)= A * B for the 32 x 32 quarter (wr, wc) of a 64 x 64 x 64 product, A and B in
// LDS with leading dimension LD.  v_mfma_f64_16x16x4 lane layout: A[m = l%16][k = l/16],
// B[k = l/16][n = l%16], D[m = 4 r + l/16][n = l%16] (register r of lane l).
__device__ inline void mfma_block(const double* A, const double* B, v4d acc[2][2], int wr, int wc,
                                  int lane, double sign) {
    const int l16 = lane & 15, l4 = lane >> 4;
#pragma unroll 4
    for (int kk = 0; kk < kNB; kk += 4) {
        double a0 = sign * A[(wr * 32 + l16) * LD + kk + l4];
        double a1 = sign * A[(wr * 32 + 16 + l16) * LD + kk + l4];
        double b0 = B[(kk + l4) * LD + wc * 32 + l16];
        double b1 = B[(kk + l4) * LD + wc * 32 + 16 + l16];
        acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
        acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
        acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
        acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
    }
}

// Stores the wave's accumulators into an LDS block.
__device__ inline void acc_to_lds(double* C, const v4d acc[2][2], int wr, int wc, int lane) {
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r)
                C[(wr * 32 + i * 16 + 4 * r + l4) * LD + wc * 32 + j * 16 + l16] = acc[i][j][r];
}

// Loads a rows x cols block of row-major global memory (leading dimension ld) into LDS,
// zero-padding to 64 x 64.  Fixed trip count (256 threads x 16) so all 16 loads of a thread are
// in flight together.
__device__ inline void load_block(double* dst, const double* src, long long ld, int rows, int cols) {
    double v[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + it * 256, r = e >> 6, c = e & 63;
        v[it] = (r < rows && c < cols) ? src[(long long)r * ld + c] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + it * 256;
        dst[(e >> 6) * LD + (e & 63)] = v[it];
    }
}

__device__ inline void store_block(double* dst, const double* src, long long ld, int rows, int cols) {
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + it * 256, r = e >> 6, c = e & 63;
        if (r < rows && c < cols) dst[(long long)r * ld + c] = src[r * LD + c];
    }
}

// Factors D (kb x kb, no interchanges) into L\U and writes L^{-1} into Li and U^{-1} into Ui
// (both zero outside the kb x kb corner).  Sets *bad if a multiplier exceeds 1 in magnitude.
//
// The elimination runs on 4 x 4 tiles: thread t owns tile (t / 16, t % 16) of the 64 x 64 block
// (rows from kb on are padded with identity rows) in registers.  Block step kk (16 of them, two
// barriers each):
//   the tiles below the diagonal tile form L_ik = A_ik U_kk^{-1}, the tiles right of it
//   U_kj = L_kk^{-1} A_kj, and publish them;
//   every trailing tile applies A_ij -= L_ik U_kj; the thread of tile (kk + 1, kk + 1) then
//   factors it (getf2 order: multipliers by the reciprocal pivot) and publishes L \ U and the
//   reciprocals for the next step (step 0's diagonal tile is factored before the loop).
// Every entry sees the same multiply-adds in the same order as in the unblocked right-looking
// elimination (step p's update of entry (r, c), for p = 0 .. min(r, c) - 1), so the factors are
// those of getf2 without interchanges.  The published tiles are stored element-major (Lp[e][i],
// Up[e][j]): the 16 threads of a row of tiles read 16 consecutive doubles of Up and 4 of Lp, free
// of bank conflicts.  A step takes about 1.2 us, most of it the dependent chain of the diagonal
// tile (four reciprocals with their Newton steps) and the two barrier-separated phases; forming
// L_ik and U_kj in every trailing tile instead (one barrier per step) made the steps slower.
// The inverses then follow in blocks (invert_factors).
__device__ inline void eliminate_tiles(double* D, int kb, int& my_bad) {
    __shared__ double Dk[16], Dinv[4];             // the diagonal tile of the step, its reciprocals
    __shared__ double Lp[16][16], Up[16][16];      // the step's L_ik (Lp[e][i]) and U_kj (Up[e][j]) tiles
    const int t = threadIdx.x, ti = t >> 4, tj = t & 15;
    double A[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gr = 4 * ti + r, gc = 4 * tj + c;
            A[r][c] = gr < kb ? D[gr * LD + gc] : (gc == gr ? 1.0 : 0.0);
        }
    // factors this thread's tile as the diagonal tile of step k and publishes it
    auto factor_diag = [&](int k) {
        double inv[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            inv[p] = fast_rcp(A[p][p]);
#pragma unroll
            for (int r = p + 1; r < 4; ++r) {
                const double l = A[r][p] * inv[p];
                if (4 * k + r < kb && !(fabs(l) <= 1.0)) my_bad = 1;
                A[r][p] = l;
#pragma unroll
                for (int c = p + 1; c < 4; ++c) A[r][c] = fma(-l, A[p][c], A[r][c]);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) Dk[4 * r + c] = A[r][c];
#pragma unroll
        for (int p = 0; p < 4; ++p) Dinv[p] = inv[p];
    };
    if (t == 0) factor_diag(0);
    __syncthreads();
#pragma unroll 1
    for (int kk = 0; kk < 16; ++kk) {
        if (tj == kk && ti > kk) {   // L_ik = A_ik U_kk^{-1}, column by column
            double U[4][4], inv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) U[r][c] = Dk[4 * r + c];
#pragma unroll
            for (int p = 0; p < 4; ++p) inv[p] = Dinv[p];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double x = A[r][c];
#pragma unroll
                    for (int p = 0; p < c; ++p) x = fma(-A[r][p], U[p][c], x);
                    const double l = x * inv[c];
                    if (4 * ti + r < kb && !(fabs(l) <= 1.0)) my_bad = 1;
                    A[r][c] = l;
                }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) Lp[4 * r + c][ti] = A[r][c];
        } else if (ti == kk && tj > kk) {   // U_kj = L_kk^{-1} A_kj, step by step
            double L[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) L[r][c] = Dk[4 * r + c];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int r = p + 1; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) A[r][c] = fma(-L[r][p], A[p][c], A[r][c]);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) Up[4 * r + c][tj] = A[r][c];
        }
        __syncthreads();
        if (ti > kk && tj > kk) {   // A_ij -= L_ik U_kj, the four steps in order
            double L[4][4], U[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    L[r][c] = Lp[4 * r + c][ti];
                    U[r][c] = Up[4 * r + c][tj];
                }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) A[r][c] = fma(-L[r][p], U[p][c], A[r][c]);
            if (ti == kk + 1 && tj == kk + 1) factor_diag(kk + 1);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) D[(4 * ti + r) * LD + 4 * tj + c] = A[r][c];
}

// One level of the blocked inversion: for every pair of H x H diagonal blocks, T = C A^{-1}
// (lower) and E B^{-1} (upper), then the off-diagonal block -B^{-1} T (lower) and -A^{-1} T (upper).
// Thread t owns the entry (t / H, t % H) of each of the pairs x 2 products; the sums run over the
// whole block (the inverses are zero outside their triangles), fully unrolled so every LDS read
// is independent of the accumulation.
template <int H>
__device__ inline void combine_level(const double* D, double* Li, double* Ui, double* T) {
    constexpr int per = H * H, pairs = kNB / (2 * H), E = pairs * 2 * per / 256;
    static_assert(per == 64 || per == 256 || per == 1024, "H is 8, 16 or 32");
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = t + 256 * q, p = e / (2 * per), lu = (e / per) & 1, idx = e % per;
        const int r = idx / H, cc = idx % H, o0 = 2 * H * p, o1 = o0 + H;
        double a0 = 0.0, a1 = 0.0;
        if (lu == 0) {
#pragma unroll
            for (int k = 0; k < H; k += 2) {
                a0 += D[(o1 + r) * LD + o0 + k] * Li[(o0 + k) * LD + o0 + cc];
                a1 += D[(o1 + r) * LD + o0 + k + 1] * Li[(o0 + k + 1) * LD + o0 + cc];
            }
        } else {
#pragma unroll
            for (int k = 0; k < H; k += 2) {
                a0 += D[(o0 + r) * LD + o1 + k] * Ui[(o1 + k) * LD + o1 + cc];
                a1 += D[(o0 + r) * LD + o1 + k + 1] * Ui[(o1 + k + 1) * LD + o1 + cc];
            }
        }
        T[e] = a0 + a1;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = t + 256 * q, p = e / (2 * per), lu = (e / per) & 1, idx = e % per;
        const int r = idx / H, cc = idx % H, o0 = 2 * H * p, o1 = o0 + H;
        const double* Tb = T + (e - idx);
        double a0 = 0.0, a1 = 0.0;
        if (lu == 0) {
#pragma unroll
            for (int k = 0; k < H; k += 2) {
                a0 += Li[(o1 + r) * LD + o1 + k] * Tb[k * H + cc];
                a1 += Li[(o1 + r) * LD + o1 + k + 1] * Tb[(k + 1) * H + cc];
            }
            Li[(o1 + r) * LD + o0 + cc] = -(a0 + a1);
        } else {
#pragma unroll
            for (int k = 0; k < H; k += 2) {
                a0 += Ui[(o0 + r) * LD + o0 + k] * Tb[k * H + cc];
                a1 += Ui[(o0 + r) * LD + o0 + k + 1] * Tb[(k + 1) * H + cc];
            }
            Ui[(o0 + r) * LD + o1 + cc] = -(a0 + a1);
        }
    }
    __syncthreads();
}

// C[m0.., n0..] (16 x 16) = sign * sum_{k < K} A[m0 + m][a0 + k] B[b0 + k][n0 + n] on one wave with
// v_mfma_f64_16x16x4 (lane layout as in mfma_block); A, B, C in LDS with the given leading dimensions.
template <int K>
__device__ inline void mfma_tile16(const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                                   double sign, int lane) {
    const int l16 = lane & 15, l4 = lane >> 4;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < K; kk += 4) acc = mfma_f64(A[l16 * lda + kk + l4], B[(kk + l4) * ldb + l16], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) C[(4 * r + l4) * ldc + l16] = sign * acc[r];
}

// combine_level for H = 16 and 32 on the matrix cores: the same two products per pair and factor,
// as 16 x 16 output tiles, one or two per wave.
template <int H>
__device__ inline void combine_level_mfma(const double* D, double* Li, double* Ui, double* T) {
    static_assert(H == 16 || H == 32, "16 x 16 tiles");
    constexpr int pairs = kNB / (2 * H), per = H * H, tpp = (H / 16) * (H / 16);   // tiles per product
    constexpr int tiles = pairs * 2 * tpp;                                           // 8 (H = 32) or 4 (H = 16)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // T = C A^{-1} (lower) and E B^{-1} (upper), T of pair p and factor lu at T + (2 p + lu) per
    for (int q = wv; q < tiles; q += 4) {
        const int p = q / (2 * tpp), lu = (q / tpp) & 1, tt = q % tpp;
        const int tr = 16 * (tt / (H / 16)), tc = 16 * (tt % (H / 16)), o0 = 2 * H * p, o1 = o0 + H;
        double* Tb = T + (2 * p + lu) * per;
        if (lu == 0)
            mfma_tile16<H>(D + (o1 + tr) * LD + o0, LD, Li + o0 * LD + o0 + tc, LD, Tb + tr * H + tc, H, 1.0, lane);
        else
            mfma_tile16<H>(D + (o0 + tr) * LD + o1, LD, Ui + o1 * LD + o1 + tc, LD, Tb + tr * H + tc, H, 1.0, lane);
    }
    __syncthreads();
    // the off-diagonal blocks: -B^{-1} T (lower) and -A^{-1} T (upper)
    for (int q = wv; q < tiles; q += 4) {
        const int p = q / (2 * tpp), lu = (q / tpp) & 1, tt = q % tpp;
        const int tr = 16 * (tt / (H / 16)), tc = 16 * (tt % (H / 16)), o0 = 2 * H * p, o1 = o0 + H;
        const double* Tb = T + (2 * p + lu) * per;
        if (lu == 0)
            mfma_tile16<H>(Li + (o1 + tr) * LD + o1, LD, Tb + tc, H, Li + (o1 + tr) * LD + o0 + tc, LD, -1.0, lane);
        else
            mfma_tile16<H>(Ui + (o0 + tr) * LD + o0, LD, Tb + tc, H, Ui + (o0 + tr) * LD + o1 + tc, LD, -1.0, lane);
    }
    __syncthreads();
}

// Inverts the unit lower factor L and the upper factor U held together in D (L\U, 64 x 64) into
// Li = L^{-1} and Ui = U^{-1}, blocked:
//   the sixteen 8 x 8 diagonal blocks (eight of L, eight of U) by substitution, one lane per
//   column, all at once;
//   then the off-diagonal blocks level by level (block sizes h = 8, 16, 32), from
//     L = [A 0; C B]  ->  L^{-1} = [A^-1 0; -B^-1 (C A^-1)  B^-1],
//     U = [A E; 0 B]  ->  U^{-1} = [A^-1  -A^-1 (E B^-1); 0 B^-1],
//   the products through LDS (T: 2048 doubles of scratch).
// Seven barriers in place of one per row: the 64-step substitution it replaces spent ~0.3 us per
// barrier-separated step.
__device__ void invert_factors(const double* D, double* Li, double* Ui, double* T) {
    const int t = threadIdx.x;
    for (int e = t; e < kNB * LD; e += 256) {
        Li[e] = 0.0;
        Ui[e] = 0.0;
    }
    __syncthreads();
    {
        const int blk = t >> 4, c = t & 15;   // 16 groups of 16 threads, 8 lanes of each work
        if (c < 8) {
            double x[8];
            if (blk < 8) {   // L block: x[i] = [i == c] - sum_{k < i} L[i][k] x[k]
                const int o = 8 * blk;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    double acc = (i == c) ? 1.0 : 0.0;
#pragma unroll
                    for (int k = 0; k < i; ++k) acc -= D[(o + i) * LD + o + k] * x[k];
                    x[i] = (i < c) ? 0.0 : acc;
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) Li[(o + i) * LD + o + c] = x[i];
            } else {         // U block: x[i] = ([i == c] - sum_{k > i} U[i][k] x[k]) / U[i][i]
                const int o = 8 * (blk - 8);
#pragma unroll
                for (int i = 7; i >= 0; --i) {
                    double acc = (i == c) ? 1.0 : 0.0;
#pragma unroll
                    for (int k = i + 1; k < 8; ++k) acc -= D[(o + i) * LD + o + k] * x[k];
                    x[i] = (i > c) ? 0.0 : acc / D[(o + i) * LD + o + i];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) Ui[(o + i) * LD + o + c] = x[i];
            }
        }
    }
    __syncthreads();
    combine_level<8>(D, Li, Ui, T);
    combine_level_mfma<16>(D, Li, Ui, T);
    combine_level_mfma<32>(D, Li, Ui, T);
}

__device__ void factor_invert_diag(double* D, double* Li, double* Ui, double* T, int kb, int* bad) {
    const int t = threadIdx.x, r = t >> 2, part = t & 3;
    int my_bad = 0;
    // ---- elimination: D becomes L (strictly below the diagonal) \ U, padded rows identity
    eliminate_tiles(D, kb, my_bad);
    if (my_bad) *bad = 1;
    __syncthreads();
    invert_factors(D, Li, Ui, T);   // ends with a barrier
    if (kb < kNB) {   // zero outside the kb x kb corner (padded rows were identity rows)
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int c = part + 4 * m;
            if (!(r < kb && c < kb)) {
                D[r * LD + c] = 0.0;
                Li[r * LD + c] = 0.0;
                Ui[r * LD + c] = 0.0;
            }
        }
        __syncthreads();
    }
}

// Per-surface views for gridDim.z batches of equal size: surface s uses A + s n^2, vectors + s n,
// uinv + s P 64^2, flags + s P, status + 2 s.
struct LuView {
    double *A, *work, *y, *x, *uinv;
    int *flags, *status;
};
__device__ inline LuView view(const LuArgs& a) {
    const int s = blockIdx.z, n = a.n, P = (n + kNB - 1) / kNB;
    LuView v;
    v.A = a.A + (long long)s * n * n;
    v.work = a.work + (long long)s * n;
    v.y = a.y + (long long)s * n;
    v.x = a.x + (long long)s * n;
    v.uinv = a.uinv + (long long)s * P * kNB * kNB;
    v.flags = a.flags + (long long)s * P;
    v.status = a.status + 2 * s;
    return v;
}

// C -= A_s B_s for the 64 x 64 tile C at (R0, C0) of the trailing matrix: A_s = the previous
// panel's L21 block (rows R0.., columns kp..kp+63), B_s = its U12 block (rows kp.., columns C0..),
// staged in LDS (As, Bs).
__device__ inline void update_tile(const LuView& v, int n, int kp, int R0, int C0, double* As, double* Bs) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int rb = min(kNB, n - R0), cb = min(kNB, n - C0);
    load_block(As, v.A + (long long)R0 * n + kp, n, rb, kNB);
    load_block(Bs, v.A + (long long)kp * n + C0, n, kNB, cb);
    double* C = v.A + (long long)R0 * n + C0;
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int row = wr * 32 + i * 16 + 4 * r + l4, col = wc * 32 + j * 16 + l16;
                acc[i][j][r] = (row < rb && col < cb) ? C[(long long)row * n + col] : 0.0;
            }
    __syncthreads();
    mfma_block(As, Bs, acc, wr, wc, lane, -1.0);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int row = wr * 32 + i * 16 + 4 * r + l4, col = wc * 32 + j * 16 + l16;
                if (row < rb && col < cb) C[(long long)row * n + col] = acc[i][j][r];
            }
}

// Loads the rows x cols block of global memory at src (leading dimension ld) into the wave's
// accumulator layout of mfma_block (zero outside the block).
__device__ inline void load_acc(v4d acc[2][2], const double* src, long long ld, int rows, int cols) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int row = wr * 32 + i * 16 + 4 * r + l4, col = wc * 32 + j * 16 + l16;
                acc[i][j][r] = (row < rows && col < cols) ? src[(long long)row * ld + col] : 0.0;
            }
}

// Panel p (columns k0 .. k0 + kb - 1) together with the previous panel's trailing update
// (kp = k0 - 64; A22 -= L21 U12 with panel p-1's blocks), one launch per panel:
//   roles 0 .. 2 nrb (panel p): each applies panel p-1's update to the diagonal block A_pp (kept in
//     LDS; the diagonal blocks are never stored) and, for roles >= 1, to its own block A_ip (row
//     roles) or A_pj (column roles) in registers, then factors A_pp = L11 U11 and inverts both
//     factors; role 0 stores U11^{-1} and y_p = L11^{-1} b_p, row roles form L21_i = A_ip U11^{-1}
//     and b_i -= L21_i y_p, column roles form U12_j = L11^{-1} A_pj;
//   roles above 2 nrb: panel p-1's update of the trailing tiles (i, j), i, j > p, that no panel
//     role owns.
// Panel p's dependency chain (the diagonal block) thereby runs beside the bulk of the previous
// panel's update rather than after it.  Every block sees the same products as with separate
// update launches: the factors are unchanged bit for bit.
__global__ __launch_bounds__(256) void lu_step_kernel(LuArgs a, int k0) {
    __shared__ double D[kNB * LD], Li[kNB * LD], Ui[kNB * LD], X[kNB * LD];
    __shared__ double T[2 * kNB * kNB / 4];   // scratch of the blocked inversion
    __shared__ double yv[kNB], wv[kNB];
    __shared__ int bad;
    const LuView v = view(a);
    const int n = a.n, t = threadIdx.x;
    const int kb = min(kNB, n - k0);
    const int rem = n - k0 - kb;
    const int nrb = (rem + kNB - 1) / kNB;
    const int role = blockIdx.x;
    const int lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
    const bool prev = k0 > 0;
    const int kp = k0 - kNB;
    if (role > 2 * nrb) {   // a trailing tile of panel p-1's update (exists only when prev)
        const int r = role - 1 - 2 * nrb;
        update_tile(v, n, kp, k0 + kb + (r / nrb) * kNB, k0 + kb + (r % nrb) * kNB, D, Li);
        return;
    }
    const bool row_role = role >= 1 && role <= nrb, col_role = role > nrb;
    const int R0 = k0 + kb + (role - 1) * kNB, rb = row_role ? min(kNB, n - R0) : 0;
    const int C0 = k0 + kb + (role - 1 - nrb) * kNB, cb = col_role ? min(kNB, n - C0) : 0;
    if (t == 0) bad = 0;
    if (k0 == 0 && role == 0) {
        const int P = (n + kNB - 1) / kNB;
        for (int i = t; i < P; i += blockDim.x) v.flags[i] = 0;
    }
    // the own block (rows x cols of A), with panel p-1's update applied below
    v4d own[2][2];
    if (row_role)
        load_acc(own, v.A + (long long)R0 * n + k0, n, rb, kb);
    else if (col_role)
        load_acc(own, v.A + (long long)k0 * n + C0, n, kb, cb);
    load_block(D, v.A + (long long)k0 * n + k0, n, kb, kb);
    if (prev) {
        load_block(Li, v.A + (long long)k0 * n + kp, n, kb, kNB);    // L21 of panel p-1, block row p
        load_block(Ui, v.A + (long long)kp * n + k0, n, kNB, kb);    // U12 of panel p-1, block column p
        if (row_role) load_block(X, v.A + (long long)R0 * n + kp, n, rb, kNB);   // L21 of panel p-1, row i
        if (col_role) load_block(X, v.A + (long long)kp * n + C0, n, kNB, cb);   // U12 of panel p-1, column j
    }
    if (t < kNB) wv[t] = (t < kb) ? v.work[k0 + t] : 0.0;
    __syncthreads();
    if (prev) {
        v4d dg[2][2];
        const int l16 = lane & 15, l4 = lane >> 4;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int r = 0; r < 4; ++r)
                    dg[i][j][r] = D[(wr * 32 + i * 16 + 4 * r + l4) * LD + wc * 32 + j * 16 + l16];
        mfma_block(Li, Ui, dg, wr, wc, lane, -1.0);                  // A_pp -= L21_p U12_p
        if (row_role) mfma_block(X, Ui, own, wr, wc, lane, -1.0);   // A_ip -= L21_i U12_p
        if (col_role) mfma_block(Li, X, own, wr, wc, lane, -1.0);   // A_pj -= L21_p U12_j
        __syncthreads();
        acc_to_lds(D, dg, wr, wc, lane);
        __syncthreads();
    }
    factor_invert_diag(D, Li, Ui, T, kb, &bad);
    // y_p = L11^{-1} b_p (4 threads per row, operands staged in registers)
    {
        const int r = t >> 2, part = t & 3;
        double lv[16], bv[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int c = part + 4 * m;
            lv[m] = Li[r * LD + c];
            bv[m] = wv[c];
        }
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 16; ++m) s += lv[m] * bv[m];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (part == 0) yv[r] = (r < kb) ? s : 0.0;
    }
    __syncthreads();
    if (role == 0) {
        double* U = v.uinv + (long long)(k0 / kNB) * kNB * kNB;
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int e = t + it * 256;
            U[e] = Ui[(e >> 6) * LD + (e & 63)];
        }
        if (t < kb) v.y[k0 + t] = yv[t];
        if (t == 0 && bad) atomicOr(&v.status[0], 1);
        return;
    }
    acc_to_lds(X, own, wr, wc, lane);   // the updated own block
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    if (row_role) {
        mfma_block(X, Ui, acc, wr, wc, lane, 1.0);        // L21 = A21 U11^{-1}
        int my_bad = 0;   // the multipliers against GEPP's pivot rule, from the accumulators
        {
            const int l16 = lane & 15, l4 = lane >> 4;
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    for (int r = 0; r < 4; ++r) {
                        const int row = wr * 32 + i * 16 + 4 * r + l4, col = wc * 32 + j * 16 + l16;
                        if (row < rb && col < kb && !(fabs(acc[i][j][r]) <= 1.0)) my_bad = 1;
                    }
        }
        __syncthreads();
        acc_to_lds(X, acc, wr, wc, lane);
        __syncthreads();
        if (my_bad) atomicOr(&v.status[0], 1);
        store_block(v.A + (long long)R0 * n + k0, X, n, rb, kb);
        {                                                 // b_i -= L21_i y_p
            const int r = t >> 2, part = t & 3;
            double xv[16], yy[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int c = part + 4 * m;
                xv[m] = X[r * LD + c];
                yy[m] = yv[c];
            }
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 16; ++m) s += xv[m] * yy[m];
            s += __shfl_xor(s, 1, 64);
            s += __shfl_xor(s, 2, 64);
            if (part == 0 && r < rb) v.work[R0 + r] -= s;
        }
    } else {
        mfma_block(Li, X, acc, wr, wc, lane, 1.0);        // U12 = L11^{-1} A12
        __syncthreads();
        acc_to_lds(X, acc, wr, wc, lane);
        __syncthreads();
        store_block(v.A + (long long)k0 * n + C0, X, n, kb, cb);
    }
    if (t == 0 && bad && role == 1) atomicOr(&v.status[0], 1);
}

// x_p = U_pp^{-1} (y_p - sum_{q>p} U_pq x_q).  Workgroup b handles block row p = P-1-b, so a
// workgroup only waits for workgroups dispatched before it; every wait is bounded.  Each U_pq
// block is staged into LDS before the wait for x_q, so only the x_q load and an LDS GEMV sit on
// the dependency chain.
__global__ __launch_bounds__(256) void lu_backsolve_kernel(LuArgs a) {
    __shared__ double Ub[kNB * LD], Ui[kNB * LD];
    __shared__ double xq[kNB], part[4][kNB], sv[kNB];
    const LuView v = view(a);
    const int n = a.n, t = threadIdx.x;
    const int P = (n + kNB - 1) / kNB;
    const int p = P - 1 - blockIdx.x;
    const int R0 = p * kNB, rb = min(kNB, n - R0);
    const int r = t & 63, chunk = t >> 6;
    const double* U = v.uinv + (long long)p * kNB * kNB;
    load_block(Ui, U, kNB, kNB, kNB);
    double acc = 0.0;
    for (int q = P - 1; q > p; --q) {
        const int C0 = q * kNB, cb = min(kNB, n - C0);
        load_block(Ub, v.A + (long long)R0 * n + C0, n, rb, cb);
        if (t == 0) {
            // relaxed polling, one acquire once the flag is seen: on gfx950 every agent-scope
            // acquire invalidates this XCD's L2, which other workgroups are reading through
            int spins = 0;
            while (__hip_atomic_load(&v.flags[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1 << 22)) {
                    atomicOr(&v.status[1], 1);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (t < kNB) {
            xq[t] = (t < cb) ? v.x[C0 + t] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 16; ++c) acc += Ub[r * LD + chunk * 16 + c] * xq[chunk * 16 + c];
        __syncthreads();
    }
    part[chunk][r] = acc;
    __syncthreads();
    if (t < kNB) sv[t] = (t < rb) ? v.y[R0 + t] - (part[0][t] + part[1][t] + part[2][t] + part[3][t]) : 0.0;
    __syncthreads();
    {   // x_p = U_pp^{-1} sv, 4 threads per row
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += Ui[r * LD + chunk * 16 + c] * sv[chunk * 16 + c];
        part[chunk][r] = s;
    }
    __syncthreads();
    if (t < rb) v.x[R0 + t] = part[0][t] + part[1][t] + part[2][t] + part[3][t];
    __syncthreads();
    if (t == 0) __hip_atomic_store(&v.flags[p], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Unblocked GEPP (first row of maximal |a|) and solve by one workgroup per surface, run only
// for surfaces whose interchange-free factorisation reported a multiplier above 1
// (status[2 s] != 0).  M holds the freshly re-assembled matrix of that surface.
__global__ __launch_bounds__(256) void gepp_fallback_kernel(const SurfDesc* sd, double* M, const double* bvec,
                                                            double* xvec, int* pivv, int* status,
                                                            int* counters) {
    const int s_ = blockIdx.x;
    if (status[2 * s_] == 0) return;
    const SurfDesc dsc = sd[s_];
    const int n = dsc.n;
    double* A = M + dsc.moff;
    const double* b = bvec + dsc.voff;
    double* x = xvec + dsc.voff;
    int* piv = pivv + dsc.voff;
    if (threadIdx.x == 0) atomicAdd(&counters[0], 1);
    __shared__ double rv[256];
    __shared__ int ri[256];
    const int t = threadIdx.x;
    for (int i = t; i < n; i += blockDim.x) x[i] = b[i];
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        double best = -1.0;
        int bi = n;
        for (int i = j + t; i < n; i += blockDim.x) {
            const double vv = fabs(A[(long long)i * n + j]);
            if (vv > best) { best = vv; bi = i; }
        }
        rv[t] = best;
        ri[t] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) {
                if (rv[t + s] > rv[t] || (rv[t + s] == rv[t] && ri[t + s] < ri[t])) {
                    rv[t] = rv[t + s];
                    ri[t] = ri[t + s];
                }
            }
            __syncthreads();
        }
        const int pr = ri[0];
        if (t == 0) piv[j] = pr;
        if (pr != j && pr < n) {
            for (int c = t; c < n; c += blockDim.x) {
                const double tmp = A[(long long)j * n + c];
                A[(long long)j * n + c] = A[(long long)pr * n + c];
                A[(long long)pr * n + c] = tmp;
            }
            if (t == 0) {
                const double tmp = x[j];
                x[j] = x[pr];
                x[pr] = tmp;
            }
        }
        __syncthreads();
        const double d = A[(long long)j * n + j];
        if (d == 0.0) {   // exactly singular column: getrf records it and moves on
            __syncthreads();
            continue;
        }
        const double xj = x[j];
        for (int i = j + 1 + t; i < n; i += blockDim.x) {
            double* Ri = A + (long long)i * n;
            const double l = Ri[j] / d;
            Ri[j] = l;
            const double* Rj = A + (long long)j * n;
            for (int c = j + 1; c < n; ++c) Ri[c] -= l * Rj[c];
            x[i] -= l * xj;
        }
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; --i) {
        if (t == 0) x[i] = x[i] / A[(long long)i * n + i];
        __syncthreads();
        const double xi = x[i];
        for (int r = t; r < i; r += blockDim.x) x[r] -= A[(long long)r * n + i] * xi;
        __syncthreads();
    }
}

}  // namespace sfh
