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

// Largest radix (log2) of one LDS pass: a thread keeps 2^R complex values in registers through R
// radix-2 stages, so a transform of 2^m points makes ceil(m / R) LDS round trips and barriers.
constexpr int kFftMaxLogRadix = 3;

// Group t of a pass over half widths h, 2h, ..., 2^{R-1} h (h = 2^lh) in each of the sub
// sub-transforms of length 2^lm: elements i0 + e h, e < 2^R; pos = i0 mod h.
__device__ inline int pass_group(int t, int lm, int R, int lh, int& pos) {
    const int q = t >> (lm - R), loc = t & ((1 << (lm - R)) - 1);
    const int grp = loc >> lh;
    pos = loc & ((1 << lh) - 1);
    return (q << lm) + (grp << (lh + R)) + pos;
}

// Decimation-in-time pass: radix-2 stages of half widths h, ..., 2^{R-1} h applied to each
// group in registers (stage st pairs e and e + 2^st, twiddle W_n^{(pos + (e mod 2^st) h) n / 2^{lh+st+1}}).
template <int R>
__device__ inline void dit_pass(double2* s, int n, int lm, int sub, int lh, int sign,
                                const double2* __restrict__ tw) {
    constexpr int G = 1 << R;
    const int h = 1 << lh, ng = sub << (lm - R);
    for (int t = threadIdx.x; t < ng; t += blockDim.x) {
        int pos;
        const int i0 = pass_group(t, lm, R, lh, pos);
        double2 x[G];
#pragma unroll
        for (int e = 0; e < G; ++e) x[e] = s[i0 + e * h];
#pragma unroll
        for (int st = 0; st < R; ++st) {
            const int half = 1 << st, stride = n >> (lh + st + 1);
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const double2 w = twid(tw, (pos + j * h) * stride, sign);
#pragma unroll
                for (int b = 0; b < G; b += 2 * half) {
                    const double2 u = x[b + j], v = cmul(x[b + j + half], w);
                    x[b + j] = cadd(u, v);
                    x[b + j + half] = csub(u, v);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < G; ++e) s[i0 + e * h] = x[e];
    }
    __syncthreads();
}

// Decimation-in-frequency pass: the mirror image, stages of half widths 2^{R-1} h, ..., h with
// the twiddle applied to the difference.
template <int R>
__device__ inline void dif_pass(double2* s, int n, int lm, int sub, int lh, int sign,
                                const double2* __restrict__ tw) {
    constexpr int G = 1 << R;
    const int h = 1 << lh, ng = sub << (lm - R);
    for (int t = threadIdx.x; t < ng; t += blockDim.x) {
        int pos;
        const int i0 = pass_group(t, lm, R, lh, pos);
        double2 x[G];
#pragma unroll
        for (int e = 0; e < G; ++e) x[e] = s[i0 + e * h];
#pragma unroll
        for (int st = R - 1; st >= 0; --st) {
            const int half = 1 << st, stride = n >> (lh + st + 1);
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const double2 w = twid(tw, (pos + j * h) * stride, sign);
#pragma unroll
                for (int b = 0; b < G; b += 2 * half) {
                    const double2 u = x[b + j], v = x[b + j + half];
                    x[b + j] = cadd(u, v);
                    x[b + j + half] = cmul(csub(u, v), w);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < G; ++e) s[i0 + e * h] = x[e];
    }
    __syncthreads();
}

// Decimation in time: input in lds_index order, output in natural order.
// tw[t] = exp(-2 pi i t / n), t < n (LDS or global); sign -1 forward, +1 inverse (unnormalised).
// The power-of-two sub-transforms run as radix-2^R passes from half width 1 upwards, then the
// radix-3 combination when n = 3 * 2^m.
__device__ inline void fft_dit(double2* s, const SurfDesc& d, int sign, const double2* __restrict__ tw) {
    const int n = d.n, lm = d.log2m, m = 1 << lm, sub = d.r3 ? 3 : 1;
    for (int lh = 0; lh < lm;) {
        const int R = min(kFftMaxLogRadix, lm - lh);
        switch (R) {
            case 4: dit_pass<4>(s, n, lm, sub, lh, sign, tw); break;
            case 3: dit_pass<3>(s, n, lm, sub, lh, sign, tw); break;
            case 2: dit_pass<2>(s, n, lm, sub, lh, sign, tw); break;
            default: dit_pass<1>(s, n, lm, sub, lh, sign, tw); break;
        }
        lh += R;
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
// fft_dit: the radix-3 split first, then radix-2^R passes from the widest half width down to 1.
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
    for (int top = lm; top > 0;) {
        const int R = min(kFftMaxLogRadix, top), lh = top - R;
        switch (R) {
            case 4: dif_pass<4>(s, n, lm, sub, lh, sign, tw); break;
            case 3: dif_pass<3>(s, n, lm, sub, lh, sign, tw); break;
            case 2: dif_pass<2>(s, n, lm, sub, lh, sign, tw); break;
            default: dif_pass<1>(s, n, lm, sub, lh, sign, tw); break;
        }
        top = lh;
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
