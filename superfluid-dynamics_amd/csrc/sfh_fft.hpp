// sfh_fft.hpp -- in-LDS FFT and the reference's spectral multiplier table, shared by the
// per-field derivative kernels and the persistent per-surface stage kernel.
//
// Sizes n = r * 2^m, r in {1, 3}.  Every function is called by all threads of a workgroup and
// strides over its work by blockDim.x; each pass ends with a workgroup barrier (a transform of
// n = 1 has no pass, so callers add their own barrier before reading the result).
#pragma once
#include "sfh_common.hpp"

namespace sfh {

__device__ inline int brev(int j, int bits) {
    return bits == 0 ? 0 : (int)(__brev((unsigned)j) >> (32 - bits));
}

// LDS slot of time/frequency sample j: (j mod r) * 2^m + bitrev_m(j / r), r in {1, 3}.
__device__ inline int lds_index(const SurfDesc& d, int j) {
    if (!d.r3) return brev(j, d.log2m);
    const int q = j % 3;
    return q * (1 << d.log2m) + brev(j / 3, d.log2m);
}

__device__ inline double2 twid(const double2* __restrict__ tw, int t, int sign) {
    double2 w = tw[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

constexpr int kFftThreads = 1024;

// Multiplication by sign * i (the quarter-turn twiddle W_{4h}^h of a fused radix-2^2 pass).
__device__ inline double2 rot_quarter(double2 a, int sign) {
    return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// Decimation in time: input in lds_index order, output in natural order.
// tw[t] = exp(-2 pi i t / n), t < n (LDS or global); sign -1 forward, +1 inverse (unnormalised).
// Consecutive radix-2 stages (half widths h and 2h) are fused into one radix-2^2 pass over quads
// (i0, i0+h, i0+2h, i0+3h): one LDS round trip and one barrier per two stages, and all index
// arithmetic is shifts and masks.
__device__ inline void fft_dit(double2* s, const SurfDesc& d, int sign, const double2* __restrict__ tw) {
    const int n = d.n, lm = d.log2m, m = 1 << lm, sub = d.r3 ? 3 : 1;
    int lh = 0;   // log2 of the half width of the next stage
    for (; lh + 2 <= lm; lh += 2) {
        const int h = 1 << lh;
        const int nq = sub << (lm - 2);                  // quads in the pass
        const int s1 = n >> (lh + 1), s2 = n >> (lh + 2); // twiddle strides for lengths 2h, 4h
        for (int t = threadIdx.x; t < nq; t += blockDim.x) {
            const int q = t >> (lm - 2), loc = t & ((m >> 2) - 1);
            const int grp = loc >> lh, pos = loc & (h - 1);
            const int i0 = (q << lm) + (grp << (lh + 2)) + pos;
            const double2 a = s[i0], b = s[i0 + h], c = s[i0 + 2 * h], e = s[i0 + 3 * h];
            const double2 w1 = twid(tw, pos * s1, sign), w2 = twid(tw, pos * s2, sign);
            const double2 bw = cmul(b, w1), ew = cmul(e, w1);
            const double2 a1 = cadd(a, bw), b1 = csub(a, bw), c1 = cadd(c, ew), e1 = csub(c, ew);
            const double2 cw = cmul(c1, w2), dw = rot_quarter(cmul(e1, w2), sign);
            s[i0] = cadd(a1, cw);
            s[i0 + 2 * h] = csub(a1, cw);
            s[i0 + h] = cadd(b1, dw);
            s[i0 + 3 * h] = csub(b1, dw);
        }
        __syncthreads();
    }
    if (lh < lm) {   // one radix-2 stage left (odd log2 m), half width h = m / 2
        const int h = 1 << lh, nb = sub << (lm - 1), st = n >> (lh + 1);
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            const int q = t >> (lm - 1), pos = t & (h - 1);
            const int i = (q << lm) + pos;
            const double2 u = s[i];
            const double2 v = cmul(s[i + h], twid(tw, pos * st, sign));
            s[i] = cadd(u, v);
            s[i + h] = csub(u, v);
        }
        __syncthreads();
    }
    if (d.r3) {   // X[k + q m] = sum_r W_n^{r k} W_3^{r q} F_r[k]
        const double c = -0.5, s3 = 0.86602540378443864676372317075293618 * sign;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const double2 f0 = s[k];
            const double2 f1 = cmul(s[m + k], twid(tw, k, sign));
            const double2 f2 = cmul(s[2 * m + k], twid(tw, 2 * k, sign));
            const double2 t = cadd(f1, f2), dd = csub(f1, f2);
            const double2 base = make_double2(f0.x + c * t.x, f0.y + c * t.y);
            const double2 rot = make_double2(-s3 * dd.y, s3 * dd.x);
            s[k] = cadd(f0, t);
            s[m + k] = cadd(base, rot);
            s[2 * m + k] = csub(base, rot);
        }
        __syncthreads();
    }
}

// Decimation in frequency: input in natural order, output in lds_index order.  Mirror image of
// fft_dit: stages with half widths m/2, m/4, ..., 1, fused in pairs (2h, h).
__device__ inline void fft_dif(double2* s, const SurfDesc& d, int sign, const double2* __restrict__ tw) {
    const int n = d.n, lm = d.log2m, m = 1 << lm, sub = d.r3 ? 3 : 1;
    if (d.r3) {   // y_q[j] = W_n^{j q} sum_r x[j + r m] W_3^{r q}
        const double c = -0.5, s3 = 0.86602540378443864676372317075293618 * sign;
        for (int j = threadIdx.x; j < m; j += blockDim.x) {
            const double2 x0 = s[j], x1 = s[m + j], x2 = s[2 * m + j];
            const double2 t = cadd(x1, x2), dd = csub(x1, x2);
            const double2 base = make_double2(x0.x + c * t.x, x0.y + c * t.y);
            const double2 rot = make_double2(-s3 * dd.y, s3 * dd.x);
            s[j] = cadd(x0, t);
            s[m + j] = cmul(cadd(base, rot), twid(tw, j, sign));
            s[2 * m + j] = cmul(csub(base, rot), twid(tw, 2 * j, sign));
        }
        __syncthreads();
    }
    int lh = lm - 2;   // fused pair (2h, h) with h = 2^lh
    for (; lh >= 0; lh -= 2) {
        const int h = 1 << lh;
        const int nq = sub << (lm - 2);
        const int s1 = n >> (lh + 1), s2 = n >> (lh + 2);
        for (int t = threadIdx.x; t < nq; t += blockDim.x) {
            const int q = t >> (lm - 2), loc = t & ((m >> 2) - 1);
            const int grp = loc >> lh, pos = loc & (h - 1);
            const int i0 = (q << lm) + (grp << (lh + 2)) + pos;
            const double2 a = s[i0], b = s[i0 + h], c = s[i0 + 2 * h], e = s[i0 + 3 * h];
            const double2 w1 = twid(tw, pos * s1, sign), w2 = twid(tw, pos * s2, sign);
            const double2 a1 = cadd(a, c), c1 = cmul(csub(a, c), w2);
            const double2 b1 = cadd(b, e), e1 = rot_quarter(cmul(csub(b, e), w2), sign);
            s[i0] = cadd(a1, b1);
            s[i0 + h] = cmul(csub(a1, b1), w1);
            s[i0 + 2 * h] = cadd(c1, e1);
            s[i0 + 3 * h] = cmul(csub(c1, e1), w1);
        }
        __syncthreads();
    }
    if (lh == -1) {   // one radix-2 stage left (odd log2 m), half width 1
        const int nb = sub << (lm - 1), st = n >> 1;
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            const int i = t << 1;
            const double2 u = s[i], v = s[i + 1];
            s[i] = cadd(u, v);
            s[i + 1] = cmul(csub(u, v), twid(tw, 0 * st, sign));
        }
        __syncthreads();
    }
}

// Twiddle source: the LDS copy (staged after the data, when it fits) or the global table.
__device__ inline const double2* stage_twiddles(double2* smem, const SurfDesc& d, const double2* tw_all,
                                                int lds_tw) {
    const double2* tw = tw_all + d.twoff;
    if (!lds_tw) return tw;
    double2* t = smem + d.n;
    for (int k = threadIdx.x; k < d.n; k += blockDim.x) t[k] = tw[k];
    return t;   // visible after the caller's __syncthreads()
}

// First-derivative multiplier (utilities.cuh:119-147), applied to coefficient i.
__device__ inline double2 first_mult(double2 c, int i, int n) {
    const double dn = (double)n;
    if (i < n / 2) return make_double2((double)(-i) * c.y / dn, (double)i * c.x / dn);
    if (i == n / 2) return make_double2(-kPi * (double)i * c.y / dn, kPi * (double)i * c.x / dn);
    if (i == n / 2 + 1) return make_double2(0.0, 0.0);
    return make_double2((double)(-(i - n)) * c.y / dn, (double)(i - n) * c.x / dn);
}

// Second-derivative multiplier (utilities.cuh:185-199).
__device__ inline double2 second_mult(double2 c, int i, int n) {
    const double m = (i <= n / 2) ? (double)(-i * i) : (double)(-(i - n) * (i - n));
    return make_double2(m * c.x / (double)n, m * c.y / (double)n);
}

}  // namespace sfh
