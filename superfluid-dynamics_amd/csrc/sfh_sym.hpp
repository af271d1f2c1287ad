// sfh_sym.hpp -- pair sweeps of the per-surface workgroup kernel that evaluate every unordered node
// pair once.
//
// The pair kernel of the dipole matrix and of the velocity is antisymmetric,
// cot((z_j - z_k)/2) = -cot((z_k - z_j)/2): in the parts of cot_parts, p and q change sign with the
// order of the pair and id does not.  In the image cotangent cot((z_k - conj z_j)/2 + i h), p changes
// sign and q and id do not.  One evaluation therefore gives the matrix entries (k, j) and (j, k) and
// the pair's contributions to the sums of both rows, about half the arithmetic of a sweep over
// ordered pairs.
//
// Geometry.  The N nodes form nb = ceil(N / 64) blocks of 64.  A tile (I, J) pairs the rows of block
// I with the columns of block J; at offset r (0..63) lane l takes row I*64 + l and column
// J*64 + (l + r) % 64, so the 64 lanes of a wave touch 64 distinct columns.  Row sums accumulate in
// registers.  Column (mirror) sums accumulate in a register that moves one lane down after every
// offset (DPP wave rotate), so at offset r lane l holds the sum of column (l + r) % 64.
//
// Work split.  Each wave owns one block row I; with nb < 16 blocks, wpb = 16 / nb waves share a block
// row and split the offsets of each of its tiles.  Every pair of blocks is visited once: phase 0 takes
// the diagonal tile (I, I) at offsets 1..32 (at offset 32 lanes l and l + 32 take the same pair, so
// that offset adds no column sums), phases d = 1 .. (nb - 1) / 2 the tiles (I, I + d), and for even
// nb a last phase the tile (I, I + nb / 2) at offsets 1..32 when I < nb / 2 and 0..31 otherwise
// (together every pair of the two blocks once).  In any phase each block receives column sums from
// one block row only, so the column sums go into an LDS slot per sub-wave without atomics; a barrier
// separates the phases, and every sum is formed in a fixed order.
//
// Stored matrix (tile-skewed): the nb x nb tiles of a surface follow each other, tile (I, J) at
// element (I nb + J) * 4096; element (row I*64 + l, column J*64 + (l + r) % 64) sits at r * 64 + l
// of its tile.  A wave's 64 stores of one offset, direct or mirrored, fall in one contiguous row of
// 64 elements, and so do the 64 loads of one offset in the product.  Rows and columns beyond N hold 0.
#pragma once
#include <type_traits>

#include "sfh_wg.hpp"

namespace sfh {

// v from lane (l + 1) % 64 (wave rotate left by one lane; every lane active).
__device__ inline double dpp_rol1(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

struct SymGeo {
    int nb;       // blocks of 64 nodes
    int wpb;      // waves per block row
    int I, sub;   // this wave's block row and its index among the wpb waves of that row
    int l;        // lane
    bool active;  // the wave takes part (waves beyond nb * wpb idle)
};

__device__ inline SymGeo sym_geo(int n) {
    SymGeo g;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);   // wave-uniform: scalar
    g.nb = (n + kTile - 1) / kTile;
    g.wpb = kBW / g.nb;
    g.I = w / g.wpb;
    g.sub = w % g.wpb;
    g.l = threadIdx.x % kTile;
    g.active = g.I < g.nb;
    return g;
}

// sym_sweep for n a multiple of 64 (FULL) or not.
template <int NA, bool FULL, class F>
__device__ inline void sym_sweep_impl(int n, double* const (&acc)[NA], F& pair) {
    const SymGeo g = sym_geo(n);
    const int t = threadIdx.x;
    const int stride = g.nb * kTile;
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a][t] = 0.0;
    __syncthreads();
    double row[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) row[a] = 0.0;
    const int k = g.I * kTile + g.l;
    const int nphase = g.nb / 2 + 1;
    for (int ph = 0; ph < nphase; ++ph) {
        if (g.active) {
            const int J = (g.I + ph) % g.nb;
            int lo, hi;
            if (ph == 0) {
                lo = 1;
                hi = 33;
            } else if (2 * ph < g.nb) {
                lo = 0;
                hi = kTile;
            } else {
                lo = g.I < g.nb / 2 ? 1 : 0;
                hi = lo + 32;
            }
            const int r0 = lo + ((hi - lo) * g.sub) / g.wpb;
            const int r1 = lo + ((hi - lo) * (g.sub + 1)) / g.wpb;
            double col[NA];
#pragma unroll
            for (int a = 0; a < NA; ++a) col[a] = 0.0;
            // offset 32 of the diagonal tile (no column sums) is the last offset of its range
            const int rm = (ph == 0 && r1 == 33) ? 32 : r1;
            for (int r = r0; r < rm; ++r) {
                const int j = J * kTile + ((g.l + r) & (kTile - 1));
                pair(k, j, J, r, std::true_type{}, std::integral_constant<bool, FULL>{}, row, col);
#pragma unroll
                for (int a = 0; a < NA; ++a) col[a] = dpp_rol1(col[a]);
            }
            if (rm < r1) {
                const int j = J * kTile + ((g.l + 32) & (kTile - 1));
                pair(k, j, J, 32, std::false_type{}, std::integral_constant<bool, FULL>{}, row, col);
#pragma unroll
                for (int a = 0; a < NA; ++a) col[a] = dpp_rol1(col[a]);
            }
            if (r1 > r0) {
                const int c = J * kTile + ((g.l + r1) & (kTile - 1));
#pragma unroll
                for (int a = 0; a < NA; ++a) acc[a][g.sub * stride + c] += col[a];
            }
        }
        __syncthreads();
    }
    if (g.active) {
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[a][g.sub * stride + k] += row[a];
    }
    __syncthreads();
}

// The sweep.  acc[a] (a < NA) are LDS arrays of kBT doubles: slot s of node i at s * nb * 64 + i.
// pair(k, j, J, r, mirror, full, row, col) is called by every lane of an active wave for row
// k = I*64 + l and column j = J*64 + (l + r) % 64; it adds the pair's contribution to row k into
// row[] and, when mirror, to row j into col[].  mirror and full are std::integral_constant<bool>:
// mirror is false only at offset 32 of the diagonal tile, and full (n a multiple of 64) says that
// k and j are always < n, so the functor checks them only when it is false.  On return the total
// of node t is sym_total(acc[a], n); the caller synchronises before reusing acc.
template <int NA, class F>
__device__ inline void sym_sweep(int n, double* const (&acc)[NA], F&& pair) {
    if ((n & (kTile - 1)) == 0)
        sym_sweep_impl<NA, true>(n, acc, pair);
    else
        sym_sweep_impl<NA, false>(n, acc, pair);
}

// The total of node t (< n) over the slots of a sweep's accumulator.
__device__ inline double sym_total(const double* acc, int n) {
    const int nb = (n + kTile - 1) / kTile, wpb = kBW / nb, t = threadIdx.x;
    double s = 0.0;
    for (int q = 0; q < wpb; ++q) s += acc[q * nb * kTile + t];
    return s;
}

// This lane's row of the sweeps: k = I*64 + l (rows >= n are padding; k < 1024 always).
__device__ inline int sym_row(int n) {
    const SymGeo g = sym_geo(n);
    return g.I * kTile + g.l;
}

// Off-diagonal entries e = M_kj and em = M_jk of M = I/2 + K (createM.cuh:43-92) from one
// evaluation of the pair cotangents: with cot((z_k - z_j)/2) = (p - i q) id,
//   M_kj = c id (y'_k p - x'_k q) [+ ci q' id'],  M_jk = c id (x'_j q - y'_j p) [+ ci q' id'],
// where (p', q', id') are the image cotangent's parts (q' and id' symmetric in k, j).  czk = c Z'_k.
__device__ inline void entry_pair(const NodeTrig& tk, const NodeTrig& tj, double2 czk, double2 zj,
                                  const PairConst& pc, double& e, double& em) {
    const CotParts cp = cot_parts(tk, tj);
    e = cp.id * (czk.y * cp.p - czk.x * cp.q);
    em = (pc.c * cp.id) * (zj.x * cp.q - zj.y * cp.p);
    if (pc.image) {
        const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
        const double im = pc.ci * (cq.q * cq.id);
        e += im;
        em += im;
    }
}

// (M v)_t evaluated matrix-free in FP64 (entry_pair and the diagonal), v in L.vs with the padding
// entries 0 (the caller has synchronised).  Returns the value on threads t < n, 0 elsewhere.
__device__ inline double mv_free_sym(int n, const PairConst& pc, BLds& L) {
    const int t = threadIdx.x;
    const int kr = sym_row(n);
    const NodeTrig tk = L.tr[kr];
    const double2 zq = L.zp[kr];
    const double2 czk = make_double2(pc.c * zq.x, pc.c * zq.y);
    const double vk = L.vs[kr];
    double* const acc[1] = {L.part};
    sym_sweep<1>(n, acc, [&](int k, int j, int, int, auto mirror, auto full, double (&row)[1], double (&col)[1]) {
        if (!decltype(full)::value && (k >= n || j >= n)) return;
        double e, em;
        entry_pair(tk, L.tr[j], czk, L.zp[j], pc, e, em);
        row[0] += e * L.vs[j];
        if constexpr (decltype(mirror)::value) col[0] += em * vk;
    });
    double w = 0.0;
    if (t < n) w = sym_total(L.part, n) + L.diag[t] * L.vs[t];
    __syncthreads();
    return w;
}

// (M v)_t for the stored tile-skewed matrix, v in L.vs with L.vs[j] = 0 for n <= j < nb*64 (the
// caller has synchronised).  Returns the value on threads t < n, 0 elsewhere.  The waves of a block
// row split its nb * 64 offset rows; partial sums meet in L.part.
template <typename T>
__device__ inline double mv_tiles(const T* __restrict__ M, int n, BLds& L) {
    const SymGeo g = sym_geo(n);
    const int t = threadIdx.x;
    double acc0 = 0.0, acc1 = 0.0;
    if (g.active) {
        const int span = g.nb * kTile;   // offset rows of the block row
        const int i0 = (span * g.sub) / g.wpb, i1 = (span * (g.sub + 1)) / g.wpb;
        const T* base = M + (long long)g.I * g.nb * kTileElems + g.l;
        constexpr int U = sizeof(T) == 8 ? 16 : 32;   // loads in flight per lane
        int i = i0;
        for (; i + U <= i1; i += U) {
            T m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) m[u] = base[(long long)(i + u) * kTile];
#pragma unroll
            for (int u = 0; u < U; u += 2) {
                const int ia = i + u, ib = i + u + 1;
                acc0 += (double)m[u] * L.vs[(ia & ~(kTile - 1)) + ((g.l + ia) & (kTile - 1))];
                acc1 += (double)m[u + 1] * L.vs[(ib & ~(kTile - 1)) + ((g.l + ib) & (kTile - 1))];
            }
        }
        for (; i < i1; ++i)
            acc0 += (double)base[(long long)i * kTile] * L.vs[(i & ~(kTile - 1)) + ((g.l + i) & (kTile - 1))];
    }
    const int stride = g.nb * kTile;
    if (g.active) L.part[g.sub * stride + g.I * kTile + g.l] = acc0 + acc1;
    __syncthreads();
    double w = 0.0;
    if (t < n)
        for (int q = 0; q < g.wpb; ++q) w += L.part[q * stride + t];
    __syncthreads();
    return w;
}

}  // namespace sfh
