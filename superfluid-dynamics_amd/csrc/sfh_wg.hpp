// sfh_wg.hpp -- building blocks of the per-surface workgroup kernels (the batched stage kernel
// and the adjoint kernel): one workgroup of 1024 threads owns one surface of N <= 1024 nodes,
// thread t owns node t, and pair loops split the columns over G = 1024 / N groups (thread t takes
// row t % N and columns j = t / N (mod G)).  Everything a surface needs besides its matrix lives
// in LDS (BLds); reductions are workgroup-local, so no workgroup ever waits for another.
#pragma once
#include "sfh_kernels.hpp"
#include "sfh_fft.hpp"
#include "sfh_physics.hpp"

namespace sfh {

constexpr int kBT = kBatchThreads;   // threads per workgroup
constexpr int kBW = kBT / 64;         // waves per workgroup
constexpr int kBK = 8;                // Krylov vectors per GMRES cycle
constexpr int kBR = kBK + 2;          // reduction slots: kBK + 1 dots and one norm

struct BLds {
    double2 fa[kBatchMaxN];
    double2 fb[kBatchMaxN];
    double2 tw[kBatchMaxN];
    NodeTrig tr[kBatchMaxN];
    double2 zp[kBatchMaxN];
    double diag[kBatchMaxN];
    double vs[kBatchMaxN];
    double part[kBT];
    double red[kBW][kBR];
    double sums[kBR];
    double H[kBK + 1][kBK];
    double cs[kBK], sn[kBK], g[kBK + 1], y[kBK];
    double col[kBK + 2];
    double scal[4];
};

// Sums each of the C slots of v over the workgroup into L.sums (all threads take part).
template <int C>
__device__ inline void bsum(const double (&v)[C], BLds& L) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) L.red[w][i] = s;
    }
    __syncthreads();
    if (t < C) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kBW; ++q) acc += L.red[q][t];
        L.sums[t] = acc;
    }
    __syncthreads();
}

__device__ inline double bsum1(double v, BLds& L) {
    const double a[1] = {v};
    bsum(a, L);
    return L.sums[0];
}

// L.sums[i] = sum over the workgroup of V[i] * w for i <= j (0 above), L.sums[K] = |w|^2 when
// with_norm.  The products are formed inside the reduction, no array of them is kept.
template <int K>
__device__ inline void bsum_dots(const double (&V)[K], double w, int j, bool with_norm, BLds& L) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double s = wave_sum((i <= j) ? V[i] * w : 0.0);
        if (lane == 0) L.red[wv][i] = s;
    }
    if (with_norm) {
        const double s = wave_sum(w * w);
        if (lane == 0) L.red[wv][K] = s;
    }
    __syncthreads();
    if (t <= K) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kBW; ++q) acc += L.red[q][t];
        L.sums[t] = acc;
    }
    __syncthreads();
}

// Register-array element i (runtime i) without dynamic register indexing.
template <int K>
__device__ inline double rget(const double (&V)[K], int i) {
    double r = V[0];
#pragma unroll
    for (int u = 1; u < K; ++u) r = (i == u) ? V[u] : r;
    return r;
}
template <int K>
__device__ inline void rset(double (&V)[K], int i, double x) {
#pragma unroll
    for (int u = 0; u < K; ++u) V[u] = (i == u) ? x : V[u];
}

// (M v)_k for the stored column-major matrix, v in L.vs (the caller has synchronised).  Returns
// the value on threads t < n, 0 elsewhere.
template <typename T>
__device__ inline double mv_stored(const T* __restrict__ M, int n, int G, BLds& L) {
    const int t = threadIdx.x;
    double acc0 = 0.0, acc1 = 0.0;
    if (t < G * n) {
        const int k = t % n, g = t / n;
        const T* col = M + k;
        int j = g;
        constexpr int U = sizeof(T) == 8 ? 8 : 16;   // loads in flight per thread
        for (; j + (U - 1) * G < n; j += U * G) {
            T m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) m[u] = col[(size_t)(j + u * G) * n];
#pragma unroll
            for (int u = 0; u < U; u += 2) {
                acc0 += (double)m[u] * L.vs[j + u * G];
                acc1 += (double)m[u + 1] * L.vs[j + (u + 1) * G];
            }
        }
        for (; j < n; j += G) acc0 += (double)col[(size_t)j * n] * L.vs[j];
    }
    L.part[t] = acc0 + acc1;
    __syncthreads();
    double w = 0.0;
    if (t < n)
        for (int q = 0; q < G; ++q) w += L.part[q * n + t];
    __syncthreads();
    return w;
}

struct PairConst {
    double c;     // coefficient of Im(Z'_k cot) (createM.cuh:52, :78)
    double ci;    // coefficient of the image term (:87-88)
    double eh, ehi;
    int image;
};

// (M v)_k evaluated matrix-free in FP64 with the same entry formula as the FP64 assembly.
__device__ inline double mv_free(int n, int G, const PairConst& pc, BLds& L) {
    const int t = threadIdx.x;
    double val = 0.0;
    if (t < G * n) {
        const int k = t % n, g = t / n;
        const NodeTrig tk = L.tr[k];
        double sx = 0.0, sy = 0.0, si = 0.0;
        for (int j = g; j < n; j += G) {
            if (j == k) continue;
            const NodeTrig tj = L.tr[j];
            const double vj = L.vs[j];
            const CotParts cp = cot_parts(tk, tj);
            const double w = vj * cp.id;
            sx += w * cp.p;
            sy -= w * cp.q;
            if (pc.image) {
                const CotParts cq = cot_image_parts(tk, tj, pc.eh, pc.ehi);
                si -= vj * (cq.q * cq.id);
            }
        }
        const double2 zk = L.zp[k];
        val = pc.c * (zk.x * sy + zk.y * sx) - pc.ci * si;
    }
    L.part[t] = val;
    __syncthreads();
    double w = 0.0;
    if (t < n) {
        for (int q = 0; q < G; ++q) w += L.part[q * n + t];
        w += L.diag[t] * L.vs[t];
    }
    __syncthreads();
    return w;
}

// Restarted GMRES on A x = rhs over the workgroup; x (thread t < n holds x_t) is updated in
// place.  x_zero: x is known to be 0 (the first residual is rhs, no product needed); r_given:
// the first residual rhs - A x is r_first (no product needed either).  Returns the number of
// Arnoldi steps; *unconv is set when tol_abs was not reached.
template <class MV>
__device__ int wg_gmres(MV&& mv, int n, double rhs, double& x, bool x_zero, double tol_abs, int cycles,
                        BLds& L, int* unconv, bool r_given = false, double r_first = 0.0) {
    const int t = threadIdx.x;
    const bool own = t < n;
    int total = 0;
    for (int cyc = 0; cyc < cycles; ++cyc) {
        double r = own ? rhs : 0.0;
        if (cyc == 0 && r_given) {
            r = own ? r_first : 0.0;   // the caller formed rhs - M x already
        } else if (!(x_zero && cyc == 0)) {
            if (own) L.vs[t] = x;
            __syncthreads();
            const double w = mv();
            r = own ? rhs - w : 0.0;
        }
        const double beta = sqrt(bsum1(r * r, L));
        if (!(beta > tol_abs)) return total;
        double V[kBK + 1];
#pragma unroll
        for (int i = 0; i <= kBK; ++i) V[i] = 0.0;
        V[0] = r / beta;
        if (t == 0) L.g[0] = beta;
        int k = 0;
        bool conv = false;
        for (int j = 0; j < kBK; ++j) {
            if (own) L.vs[t] = rget(V, j);
            __syncthreads();
            double w = mv();
            // classical Gram-Schmidt ...
            bsum_dots(V, w, j, false, L);
            if (t == 0)
                for (int i = 0; i <= j; ++i) L.col[i] = L.sums[i];
#pragma unroll
            for (int i = 0; i <= kBK; ++i)
                if (i <= j) w -= L.sums[i] * V[i];
            // ... with one re-orthogonalisation; |w|^2 after it by Pythagoras (the second-pass
            // coefficients are O(eps |w|), so the subtraction does not cancel)
            __syncthreads();
            bsum_dots(V, w, j, true, L);
            double hh = L.sums[kBK + 1];
#pragma unroll
            for (int i = 0; i <= kBK; ++i) {
                if (i <= j) {
                    const double c2 = L.sums[i];
                    w -= c2 * V[i];
                    hh -= c2 * c2;
                }
            }
            const double hn = sqrt(fmax(hh, 0.0));
            rset(V, j + 1, (own && hn > 0.0) ? w / hn : 0.0);
            if (t == 0) {   // new Hessenberg column, previous rotations, new rotation
                double* col = L.col;
                for (int i = 0; i <= j; ++i) col[i] += L.sums[i];
                col[j + 1] = hn;
                for (int i = 0; i < j; ++i) {
                    const double a0 = L.cs[i] * col[i] + L.sn[i] * col[i + 1];
                    const double a1 = -L.sn[i] * col[i] + L.cs[i] * col[i + 1];
                    col[i] = a0;
                    col[i + 1] = a1;
                }
                const double den = hypot(col[j], col[j + 1]);
                const double cj = den > 0.0 ? col[j] / den : 1.0;
                const double sj = den > 0.0 ? col[j + 1] / den : 0.0;
                L.cs[j] = cj;
                L.sn[j] = sj;
                col[j] = den;
                for (int i = 0; i <= j; ++i) L.H[i][j] = col[i];
                L.g[j + 1] = -sj * L.g[j];
                L.g[j] = cj * L.g[j];
                L.scal[0] = fabs(L.g[j + 1]);
                L.scal[1] = hn;
            }
            __syncthreads();
            k = j + 1;
            if (L.scal[0] <= tol_abs || L.scal[1] == 0.0) {
                conv = L.scal[0] <= tol_abs;
                break;
            }
        }
        // x += V_k y with H_k y = g (upper triangular after the rotations)
        if (t == 0) {
            for (int i = k - 1; i >= 0; --i) {
                double s = L.g[i];
                for (int l = i + 1; l < k; ++l) s -= L.H[i][l] * L.y[l];
                L.y[i] = s / L.H[i][i];
            }
        }
        __syncthreads();
        if (own) {
#pragma unroll
            for (int i = 0; i < kBK; ++i)
                if (i < k) x += L.y[i] * V[i];
        }
        __syncthreads();
        total += k;
        if (conv) return total;
    }
    if (unconv) *unconv = 1;
    return total;
}

// Sum of v over the G column groups of row t % n (inactive threads pass 0); valid on t < n.
__device__ inline double gsum(double v, int n, int G, BLds& L) {
    const int t = threadIdx.x;
    L.part[t] = v;
    __syncthreads();
    double w = 0.0;
    if (t < n)
        for (int q = 0; q < G; ++q) w += L.part[q * n + t];
    __syncthreads();
    return w;
}

// Node values produced by the derivative phase (valid on threads t < n).
struct NodeVals {
    double2 z, zp, zpp;
    double phi, b;
};

// Spectral derivatives of one surface in LDS (Derivatives.cuh:310-384): Z' and Z'' from ONE
// forward transform of Z - 2 pi j / N, Re Phi' from a second; also fills L.tw (twiddles),
// L.tr (half-angle table) and L.zp.
__device__ inline NodeVals wg_derivatives(const SurfDesc& d, const double2* __restrict__ tw_all,
                                          const double2* __restrict__ Zin, const double* __restrict__ Pin,
                                          const Phys& ph, BLds& L) {
    const int n = d.n, off = d.voff, t = threadIdx.x;
    const bool own = t < n;
    const double twopi_n = 2.0 * kPi / n;
    for (int i = t; i < n; i += kBT) L.tw[i] = tw_all[d.twoff + i];
    NodeVals v;
    v.z = make_double2(0.0, 0.0);
    v.phi = 0.0;
    if (own) {
        v.z = Zin[off + t];
        v.phi = Pin[off + t];
        L.tr[t] = node_trig(v.z);
        L.fa[lds_index(d, t)] = make_double2(v.z.x - twopi_n * (double)t, v.z.y);   // :285,317
    }
    __syncthreads();
    fft_dit(L.fa, d, -1, L.tw);
    __syncthreads();
    for (int i = t; i < n; i += kBT) {
        const double2 c = L.fa[i];
        L.fb[i] = second_mult(c, i, n);
        L.fa[i] = first_mult(c, i, n);
    }
    __syncthreads();
    fft_dif(L.fa, d, +1, L.tw);
    fft_dif(L.fb, d, +1, L.tw);
    __syncthreads();
    v.zp = make_double2(1.0, 0.0);
    v.zpp = make_double2(0.0, 0.0);
    if (own) {
        const double2 a = L.fa[lds_index(d, t)], c = L.fb[lds_index(d, t)];
        const double sc2 = 4.0 * kPi * kPi / ((double)n * n);
        v.zp = make_double2(a.x * twopi_n + twopi_n, a.y * twopi_n);   // :323, :374
        v.zpp = make_double2(c.x * sc2, c.y * sc2);                    // :321
        L.zp[t] = v.zp;
    }
    __syncthreads();
    const double plin = -(1 + ph.rho) * kPi * ph.U / n;   // Derivatives.cuh:286
    if (own) L.fa[lds_index(d, t)] = make_double2(v.phi - plin * (double)t, 0.0);
    __syncthreads();
    fft_dit(L.fa, d, -1, L.tw);
    __syncthreads();
    for (int i = t; i < n; i += kBT) L.fa[i] = first_mult(L.fa[i], i, n);
    __syncthreads();
    fft_dif(L.fa, d, +1, L.tw);
    __syncthreads();
    v.b = 0.0;
    if (own) {
        v.b = L.fa[lds_index(d, t)].x * twopi_n;   // :324, BaseBoundaryIntegrator.cuh:299
        if (ph.U != 0) v.b += plin;                 // :380-383
    }
    __syncthreads();
    return v;
}

// Pair constants and the diagonal M_kk (createM.cuh:43-92) into L.diag.
__device__ inline PairConst wg_diagonal(const Phys& ph, const NodeVals& v, int n, BLds& L) {
    const bool water = ph.kind == 0;
    PairConst pc;
    pc.c = water ? 0.25 * (1 - ph.rho) / kPi : 0.25 / kPi;
    pc.ci = 0.25 / kPi;
    pc.eh = exp(ph.depth);
    pc.ehi = exp(-ph.depth);
    pc.image = ph.image;
    if ((int)threadIdx.x < n) {
        const double2 r = cdiv(v.zpp, v.zp);
        double dg = (water ? 0.5 * (1 + ph.rho) : 0.5) + pc.c * r.y;
        if (ph.image) dg -= pc.ci * cot64(0.0, v.z.y + ph.depth).y;
        L.diag[threadIdx.x] = dg;
    }
    __syncthreads();
    return pc;
}

// In-LDS spectral first derivative of a field already stored at L.fa[lds_index(j)]; the result is
// left at the same slots (unscaled).  conj_mult applies the adjoint table (every first-derivative
// multiplier is imaginary, so the adjoint is the negated table).
__device__ inline void wg_first_derivative(const SurfDesc& d, BLds& L, bool adjoint) {
    const int n = d.n, t = threadIdx.x;
    __syncthreads();
    fft_dit(L.fa, d, -1, L.tw);
    __syncthreads();
    for (int i = t; i < n; i += kBT) {
        const double2 c = first_mult(L.fa[i], i, n);
        L.fa[i] = adjoint ? make_double2(-c.x, -c.y) : c;
    }
    __syncthreads();
    fft_dif(L.fa, d, +1, L.tw);
    __syncthreads();
}

}  // namespace sfh
