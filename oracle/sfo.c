/*
 * sfo.c -- CPU FP64 oracle for the Roberts (1983) periodic boundary-integral RK4 stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP implementation: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load it.  The product
 * library never links or calls it.
 *
 * It restates, function by function, the FP64 algorithm of the reference CUDA solver
 * (paths relative to CuSuperHelium/CuSuperHelium/):
 *   sfo_cot               utilities.cuh:312-315,347-350      cot(w) = 1/tan(w)
 *   sfo_fft               cufftExecZ2Z (Derivatives.cuh:197,243), unnormalised both ways
 *   sfo_fft_derivative    Derivatives.cuh:189-257 with the multiplier tables
 *                         utilities.cuh:106-148 (first) and 178-200 (second)
 *   sfo_zphi_derivative   Derivatives.cuh:274-384
 *   sfo_assemble          createM.cuh:43-63 (water) and 65-92 (finite depth / helium)
 *   sfo_lu_factor/solve   cusolverDnDgetrf/getrs as called in MatrixSolver.cuh:113-126
 *   sfo_velocity          WaterVelocities.cuh:38-115,205-242
 *   sfo_rhs_phi           createM.cuh:96-117,171-211
 *   sfo_rhs               BaseBoundaryIntegrator.cuh:137-306 (energies omitted: they do not
 *                         feed the right-hand side)
 *   sfo_rk4               AutonomousRungeKuttaStepper.cuh:31-36,123-361,417-437
 *   sfo_intensity         LightIntensity::compute_intensity (LightIntensity.cuh:16-29)
 *   sfo_rhs_driven        HeliumDrivenAutonomousProblem::CalculateRhsPhi (HeliumDrivenAutonomousProblem.cuh:19-25,
 *                         add_optical_field_drive_terms_no_time_depence createM.cuh:138-149) and the
 *                         augmented system's DelayedIntensityIntegrator::run (DelayedIntensityIntegrator.cuh:20-32,
 *                         createM.cuh:151-169; AugmentedBoundaryIntegrator.cuh:32-36)
 *   sfo_rk4_driven        the same RK4 over [Z, Phi(, D)]
 *   sfo_rk4_timed         the time-dependent drive and its RK4 (RK4_Time_Dependent.cuh:170-412,
 *                         createM.cuh:119-136, DelayedIntensityTerm.cuh:14-38)
 *   sfo_jacobian          JacobianCalculator::calculateJacobian (JacobianCalculator.cuh:223-284): central
 *                         differences of the RHS, column-major 3N x 3N
 *   sfo_gl2               GaussLegendre2::runEvolution / gaussLegendreS2Step (GaussLegendre.cuh:190-560),
 *                         two-stage Gauss-Legendre with Newton, Armijo backtracking and step halving
 *   sfo_energies          Energies.cuh:56-204 (kinetic, gravitational / van der Waals, surface,
 *                         volume flux), called per RHS at BaseBoundaryIntegrator.cuh:270-272
 *   sfo_adimensionalize   Export.cu:1222-1246
 *
 * Layouts follow the reference: a batch of B surfaces of N nodes is the complex128 state
 * [Z_0..Z_{BN-1}, Phi_0..Phi_{BN-1}] (BaseBoundaryIntegrator.cuh:141-146), stored here as
 * interleaved doubles.  Matrices are row-major per surface: M[b*N*N + k*N + j] is row k
 * (collocation node) and column j (source node), i.e. element (k, j) of the system the
 * reference hands to getrf (createM.cuh:52 writes it at k + j*n in column-major order).
 *
 * Independence from the HIP path: cot uses glibc's ctan, the FFT is a plain iterative radix-2
 * (or recursive mixed-radix) transform, and the solve is classic LAPACK-style blocked GEPP
 * with explicit row interchanges.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double PI_D = 3.14159265358979323846264338327950288;

/* Problem description.  kind: 0 = water (WaterBoundaryProblem.cuh:7-40),
 * 1 = helium film (HeliumBoundaryProblem.cuh:45-86).  g multiplies the restoring term of the
 * Bernoulli equation; g = 1 is the reference (which has no such parameter). */
typedef struct sfo_props {
    int kind;
    int infinite_depth;
    int use_expansions;
    int expansion_order;
    int rho_quirk;       /* 1: water Bernoulli reads V1[1] as in createM.cuh:104 */
    int image_zp;        /* test-only, 0 = the reference: 1 multiplies the film image term of M by
                            Z'_k like the direct term (createM.cuh:80,87-88 omit that factor) */
    double rho;
    double kappa;
    double depth;
    double U;
    double g;
} sfo_props;

int sfo_abi_version(void) { return 3; }

int sfo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void sfo_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ cot */
/* utilities.cuh:312-315: cot(z) = 1.0 / tan(z) (libcu++ complex tan). */
static inline cplx cot_c(cplx z) { return 1.0 / ctan(z); }

void sfo_cot(double re, double im, double *out)
{
    cplx c = cot_c(re + I * im);
    out[0] = creal(c);
    out[1] = cimag(c);
}

/* ------------------------------------------------------------------ FFT */
/* Unnormalised DFT X_k = sum_j x_j exp(sign * 2 pi i j k / n), sign = -1 forward (cuFFT
 * CUFFT_FORWARD), +1 inverse (CUFFT_INVERSE, no 1/n). */
static cplx twiddle(long long num, long long den, int sign)
{
    long long r = num % den;
    if (r < 0) r += den;
    double ang = 2.0 * PI_D * (double)r / (double)den;
    return cos(ang) + I * (sign * sin(ang));
}

static void fft_pow2(cplx *a, int n, int sign)
{
    /* bit reversal */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            cplx w = twiddle(k, len, sign);
            for (int i = k; i < n; i += len) {
                cplx u = a[i], v = a[i + half] * w;
                a[i] = u + v;
                a[i + half] = u - v;
            }
        }
    }
}

static void fft_generic(cplx *a, int n, int sign, cplx *scratch)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) { fft_pow2(a, n, sign); return; }
    if (n % 2 == 0) {
        int h = n / 2;
        cplx *ev = scratch, *od = scratch + h;
        for (int i = 0; i < h; ++i) { ev[i] = a[2 * i]; od[i] = a[2 * i + 1]; }
        cplx *sub = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        fft_generic(ev, h, sign, sub);
        fft_generic(od, h, sign, sub);
        free(sub);
        for (int k = 0; k < h; ++k) {
            cplx t = twiddle(k, n, sign) * od[k];
            a[k] = ev[k] + t;
            a[k + h] = ev[k] - t;
        }
        return;
    }
    /* odd length: direct DFT */
    for (int k = 0; k < n; ++k) {
        cplx s = 0;
        for (int j = 0; j < n; ++j) s += a[j] * twiddle((long long)j * k, n, sign);
        scratch[k] = s;
    }
    memcpy(a, scratch, sizeof(cplx) * (size_t)n);
}

void sfo_fft(double *data, int n, int sign)
{
    cplx *a = (cplx *)data;
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    fft_generic(a, n, sign, scratch);
    free(scratch);
}

/* ------------------------------------------------------ spectral derivatives */
/* first_derivative_multiplication, utilities.cuh:106-148 (quirks: Nyquist times i*pi*(n/2)/n,
 * mode n/2+1 zeroed). */
static void first_mult(cplx *c, int n)
{
    for (int i = 0; i < n; ++i) {
        double x = creal(c[i]), y = cimag(c[i]);
        double rx, ry;
        if (i < n / 2) {
            rx = (double)(-i) * y / (double)n;
            ry = (double)i * x / (double)n;
        } else if (i == n / 2) {
            rx = -PI_D * (double)i * y / (double)n;
            ry = PI_D * (double)i * x / (double)n;
        } else if (i == n / 2 + 1) {
            rx = 0.0;
            ry = 0.0;
        } else {
            rx = (double)(-(i - n)) * y / (double)n;
            ry = (double)(i - n) * x / (double)n;
        }
        c[i] = rx + I * ry;
    }
}

/* second_derivative_fft, utilities.cuh:178-200 */
static void second_mult(cplx *c, int n)
{
    for (int i = 0; i < n; ++i) {
        double x = creal(c[i]), y = cimag(c[i]);
        double m = (i <= n / 2) ? (double)(-i * i) : (double)(-(i - n) * (i - n));
        c[i] = (m * x / (double)n) + I * (m * y / (double)n);
    }
}

/* FftDerivative::exec, Derivatives.cuh:189-257 (no filter). */
void sfo_fft_derivative(const double *in, double *out, int n, int second, double scale)
{
    cplx *c = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    memcpy(c, in, sizeof(cplx) * (size_t)n);
    sfo_fft((double *)c, n, -1);
    if (second) second_mult(c, n); else first_mult(c, n);
    sfo_fft((double *)c, n, +1);
    if (scale != 1.0)
        for (int i = 0; i < n; ++i) c[i] *= scale;
    memcpy(out, c, sizeof(cplx) * (size_t)n);
    free(c);
}

/* ZPhiDerivative::exec, Derivatives.cuh:274-384 for a batch of surfaces.
 * Z, Phi: complex (interleaved) B*n.  Outputs Zp, Zpp, Phip complex B*n. */
void sfo_zphi_derivative(const sfo_props *p, int n, int batch, const double *Z, const double *Phi,
                         double *Zp, double *Zpp, double *Phip)
{
    const double U = p->U, rho = p->rho;
    cplx *per = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int b = 0; b < batch; ++b) {
        const cplx *z = (const cplx *)Z + (size_t)b * n;
        const cplx *ph = (const cplx *)Phi + (size_t)b * n;
        for (int j = 0; j < n; ++j) per[j] = z[j] - 2 * PI_D * (double)j / n;      /* :285,317 */
        sfo_fft_derivative((double *)per, Zpp + 2 * (size_t)b * n, n, 1, 4.0 * PI_D * PI_D / (n * n));
        sfo_fft_derivative((double *)per, Zp + 2 * (size_t)b * n, n, 0, 2.0 * PI_D / n);
        for (int j = 0; j < n; ++j) per[j] = ph[j] - (-(1 + rho) * PI_D * U / n * (double)j); /* :286 */
        sfo_fft_derivative((double *)per, Phip + 2 * (size_t)b * n, n, 0, 2.0 * PI_D / n);
        cplx *zp = (cplx *)Zp + (size_t)b * n;
        cplx *pp = (cplx *)Phip + (size_t)b * n;
        for (int j = 0; j < n; ++j) zp[j] += 2.0 * PI_D / n;                         /* :374 */
        if (U != 0)
            for (int j = 0; j < n; ++j) pp[j] += -(1 + rho) * PI_D * U / n;          /* :380-383 */
    }
    free(per);
}

/* --------------------------------------------------------------- assembly */
/* createMKernel (createM.cuh:43-63) for kind 0, createFiniteDepthMKernel (:65-92) for kind 1.
 * With image_zp set (a test-only option, never the reference) the image term carries the factor
 * Z'_k of the direct term; with it the tangential condition Re(conj(q_k) Z'_k) = Re(Phi'_k) holds
 * for the film as it does for deep water, which pins the image terms of M and of the velocity. */
void sfo_assemble(const sfo_props *p, int n, int batch, const double *Z, const double *Zp,
                  const double *Zpp, double *M)
{
    const double rho = p->rho, h = p->depth;
    for (int b = 0; b < batch; ++b) {
        const cplx *z = (const cplx *)Z + (size_t)b * n;
        const cplx *zp = (const cplx *)Zp + (size_t)b * n;
        const cplx *zpp = (const cplx *)Zpp + (size_t)b * n;
        double *m = M + (size_t)b * n * n;
        #pragma omp parallel for schedule(static) if (n >= 256)
        for (int k = 0; k < n; ++k) {
            for (int j = 0; j < n; ++j) {
                double v;
                if (p->kind == 0) {
                    if (k == j)
                        v = 0.5 * (1 + rho) + 0.25 * (1 - rho) / PI_D * cimag(zpp[k] / zp[k]);
                    else
                        v = 0.25 * (1 - rho) / PI_D * cimag(zp[k] * cot_c(0.5 * (z[k] - z[j])));
                } else {
                    const cplx f = p->image_zp ? zp[k] : 1.0;
                    if (k == j) {
                        v = 0.5 + 0.25 / PI_D * cimag(zpp[k] / zp[k]);
                        if (!p->infinite_depth)
                            v -= 0.25 / PI_D * cimag(f * cot_c(I * (cimag(z[k]) + h)));
                    } else {
                        v = 0.25 / PI_D * cimag(zp[k] * cot_c(0.5 * (z[k] - z[j])));
                        if (!p->infinite_depth)
                            v -= 0.25 / PI_D * cimag(f * cot_c(0.5 * (z[k] - conj(z[j])) + I * h));
                    }
                }
                m[(size_t)k * n + j] = v;
            }
        }
    }
}

/* ------------------------------------------------------------------ GEPP */
/* LU with partial pivoting (row interchanges, first index of max |a|), as LAPACK dgetrf /
 * cusolverDnDgetrf.  Row-major, in place, right-looking blocked (panel width 64).
 * piv[j] = row swapped with row j at step j (0-based).  Returns 0, or j+1 for the first
 * exactly-zero pivot (the factorisation continues, as getrf does). */
int sfo_lu_factor(int n, double *A, int *piv)
{
    const int nb = 64;
    int info = 0;
    for (int j0 = 0; j0 < n; j0 += nb) {
        const int jb = (n - j0 < nb) ? n - j0 : nb;
        const int je = j0 + jb;
        for (int j = j0; j < je; ++j) {
            int pr = j;
            double amax = fabs(A[(size_t)j * n + j]);
            for (int i = j + 1; i < n; ++i) {
                double v = fabs(A[(size_t)i * n + j]);
                if (v > amax) { amax = v; pr = i; }
            }
            piv[j] = pr;
            if (pr != j) {
                double *r1 = A + (size_t)j * n, *r2 = A + (size_t)pr * n;
                for (int c = 0; c < n; ++c) { double t = r1[c]; r1[c] = r2[c]; r2[c] = t; }
            }
            const double d = A[(size_t)j * n + j];
            if (d == 0.0) { if (!info) info = j + 1; continue; }
            const double *rj = A + (size_t)j * n;
            #pragma omp parallel for schedule(static) if (n - j > 256)
            for (int i = j + 1; i < n; ++i) {
                double *ri = A + (size_t)i * n;
                double l = ri[j] / d;
                ri[j] = l;
                for (int c = j + 1; c < je; ++c) ri[c] -= l * rj[c];
            }
        }
        if (je >= n) break;
        /* U12 = L11^{-1} A12 (unit lower triangular forward substitution) */
        for (int r = j0 + 1; r < je; ++r) {
            double *rr = A + (size_t)r * n;
            for (int q = j0; q < r; ++q) {
                const double l = rr[q];
                const double *rq = A + (size_t)q * n;
                for (int c = je; c < n; ++c) rr[c] -= l * rq[c];
            }
        }
        /* A22 -= L21 U12 */
        #pragma omp parallel for schedule(static) if (n - je > 256)
        for (int i = je; i < n; ++i) {
            double *ri = A + (size_t)i * n;
            for (int q = j0; q < je; ++q) {
                const double l = ri[q];
                const double *rq = A + (size_t)q * n;
                for (int c = je; c < n; ++c) ri[c] -= l * rq[c];
            }
        }
    }
    return info;
}

/* getrs with CUBLAS_OP_N: apply the interchanges, then L y = Pb (unit), U x = y. */
void sfo_lu_solve(int n, const double *LU, const int *piv, double *b)
{
    for (int j = 0; j < n; ++j)
        if (piv[j] != j) { double t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        const double *ri = LU + (size_t)i * n;
        for (int q = 0; q < i; ++q) s -= ri[q] * b[q];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        const double *ri = LU + (size_t)i * n;
        for (int q = i + 1; q < n; ++q) s -= ri[q] * b[q];
        b[i] = s / ri[i];
    }
}

/* --------------------------------------------------------------- velocity */
/* createVelocityMatrices (WaterVelocities.cuh:38-70) / createHeliumVelocityMatrices (:72-107)
 * followed by VelocityCalculator::calculateVelocities (:205-242):
 *   w = V2 o mu' + V1 mu,  q = conj(w).
 * mu is real (the reference stores it complex with zero imaginary part), mup complex. */
void sfo_velocity(const sfo_props *p, int n, int batch, const double *Z, const double *Zp,
                  const double *Zpp, const double *mu, const double *mup, int lower, double *q)
{
    const double h = p->depth;
    const int image = (p->kind == 1) && !p->infinite_depth;
    for (int b = 0; b < batch; ++b) {
        const cplx *z = (const cplx *)Z + (size_t)b * n;
        const cplx *zp = (const cplx *)Zp + (size_t)b * n;
        const cplx *zpp = (const cplx *)Zpp + (size_t)b * n;
        const double *a = mu + (size_t)b * n;
        const cplx *ap = (const cplx *)mup + (size_t)b * n;
        cplx *out = (cplx *)q + (size_t)b * n;
        #pragma omp parallel for schedule(static) if (n >= 256)
        for (int k = 0; k < n; ++k) {
            cplx v2 = I * (1.0 / (2.0 * PI_D * zp[k]));
            cplx w = v2 * ap[k];
            for (int j = 0; j < n; ++j) {
                cplx v1;
                if (k == j) {
                    v1 = I * (-1.0 / (4.0 * PI_D) * zpp[k] / (zp[k] * zp[k]));
                    if (image) v1 += I * (1.0 / (4.0 * PI_D) * cot_c(I * (cimag(z[k]) + h)));
                    if (lower) v1 += 1.0 / (2.0 * zp[k]);
                    else v1 -= 0.5 / zp[k];
                } else {
                    v1 = I * (-1.0 / (4.0 * PI_D) * cot_c(0.5 * (z[k] - z[j])));
                    if (image) v1 += I * (1.0 / (4.0 * PI_D) * cot_c(0.5 * (z[k] - conj(z[j])) + I * h));
                }
                w += v1 * a[j];
            }
            out[k] = conj(w);
        }
    }
}

/* ------------------------------------------------------------- Bernoulli */
/* compute_rhs_phi_expression (createM.cuh:96-107), helium variants (:109-117, 171-193,
 * 200-211).  total = B*n nodes; outputs real dPhi/dt. */
void sfo_rhs_phi(const sfo_props *p, int total, const double *Z, const double *Zp, const double *Zpp,
                 const double *qlow, const double *qup, double *out)
{
    const cplx *z = (const cplx *)Z;
    const cplx *v1 = (const cplx *)qlow;
    const cplx *v2 = (const cplx *)qup;
    const double rho = p->rho, h = p->depth, g = p->g;
    for (int i = 0; i < total; ++i) {
        double y = cimag(z[i]);
        double r;
        if (p->kind == 0) {
            double a1 = creal(v1[i]) * creal(v1[i]) + cimag(v1[i]) * cimag(v1[i]);
            double a2 = creal(v2[i]) * creal(v2[i]) + cimag(v2[i]) * cimag(v2[i]);
            int qi = (p->rho_quirk && total > 1) ? 1 : i;
            double dot = creal(v1[qi]) * creal(v2[i]) + cimag(v1[i]) * cimag(v2[i]);
            r = -(1 + rho) * g * y + 0.5 * a1 + 0.5 * rho * a2 - rho * dot;
        } else if (p->use_expansions) {
            double kin = 0.5 * creal(v1[i]) * creal(v1[i]) + 0.5 * cimag(v1[i]) * cimag(v1[i]);
            double vdw = 0.0;
            switch (p->expansion_order) {
            case 3: vdw += -10.0 / 3.0 * pow(y, 3.0) / (h * h); /* fall through */
            case 2: vdw += 2.0 * pow(y, 2.0) / h;               /* fall through */
            case 1: vdw += -y;                                   /* fall through */
            default: break;
            }
            r = g * vdw + kin;
        } else if (p->kappa != 0.0) {
            const cplx *zp = (const cplx *)Zp;
            const cplx *zpp = (const cplx *)Zpp;
            double a1 = creal(v1[i]) * creal(v1[i]) + cimag(v1[i]) * cimag(v1[i]);
            double curv = (creal(zp[i]) * cimag(zpp[i]) - cimag(zp[i]) * creal(zpp[i])) /
                          pow(creal(zp[i]) * creal(zp[i]) + cimag(zp[i]) * cimag(zp[i]), 1.5);
            r = g * (20.447761896665416 * h / 3.0 * (1.0 / pow(1.0 + y / h, 3) - 1)) + 0.5 * a1 + p->kappa * curv;
        } else {
            double vdw = g * (h / 3.0);
            r = vdw * pow(1.0 + y / h, -3.0) - vdw + 0.5 * creal(v1[i]) * creal(v1[i]) +
                0.5 * cimag(v1[i]) * cimag(v1[i]);
        }
        out[i] = r;
    }
}

/* -------------------------------------------------------------------- RHS */
typedef struct sfo_diag {
    double *mu;     /* B*n real */
    double *Zp;     /* B*n complex */
    double *Zpp;    /* B*n complex */
    double *Phip;   /* B*n complex */
    double *mup;    /* B*n complex */
    double *qup;    /* B*n complex */
} sfo_diag;

/* BaseBoundaryIntegralCalculator::runTimeStep (BaseBoundaryIntegrator.cuh:137-273) and
 * calculateVorticities (:275-306).  state/rhs: complex [Z (B n), Phi (B n)] interleaved. */
int sfo_rhs(const sfo_props *p, int n, int batch, const double *state, double *rhs, sfo_diag *diag)
{
    const size_t tot = (size_t)n * batch;
    const double *Z = state;
    const double *Phi = state + 2 * tot;
    double *Zp = (double *)malloc(sizeof(double) * 2 * tot);
    double *Zpp = (double *)malloc(sizeof(double) * 2 * tot);
    double *Phip = (double *)malloc(sizeof(double) * 2 * tot);
    double *mu = (double *)malloc(sizeof(double) * tot);
    double *mucx = (double *)malloc(sizeof(double) * 2 * tot);
    double *mup = (double *)malloc(sizeof(double) * 2 * tot);
    double *qup = (double *)malloc(sizeof(double) * 2 * tot);
    double *phirhs = (double *)malloc(sizeof(double) * tot);
    double *M = (double *)malloc(sizeof(double) * (size_t)n * n);
    int *piv = (int *)malloc(sizeof(int) * (size_t)n);
    int info = 0;

    sfo_zphi_derivative(p, n, batch, Z, Phi, Zp, Zpp, Phip);
    for (int b = 0; b < batch; ++b) {
        /* CreateMMatrix + complex_to_real + MatrixSolver::solve, one surface at a time */
        sfo_assemble(p, n, 1, Z + 2 * (size_t)b * n, Zp + 2 * (size_t)b * n, Zpp + 2 * (size_t)b * n, M);
        double *a = mu + (size_t)b * n;
        for (int j = 0; j < n; ++j) a[j] = Phip[2 * ((size_t)b * n + j)];
        int inf = sfo_lu_factor(n, M, piv);
        if (inf && !info) info = inf;
        sfo_lu_solve(n, M, piv, a);
    }
    /* real_to_complex + FftDerivative::exec(a) (:201-203) */
    for (size_t i = 0; i < tot; ++i) { mucx[2 * i] = mu[i]; mucx[2 * i + 1] = 0.0; }
    for (int b = 0; b < batch; ++b)
        sfo_fft_derivative(mucx + 2 * (size_t)b * n, mup + 2 * (size_t)b * n, n, 0, 2.0 * PI_D / n);
    /* CalculateVelocities lower into rhs[0:tot], upper into scratch (:240-242) */
    sfo_velocity(p, n, batch, Z, Zp, Zpp, mu, mup, 1, rhs);
    int need_upper = (p->kind == 0 && p->rho != 0.0);
    if (need_upper || (diag && diag->qup))
        sfo_velocity(p, n, batch, Z, Zp, Zpp, mu, mup, 0, qup);
    else
        memset(qup, 0, sizeof(double) * 2 * tot);
    sfo_rhs_phi(p, (int)tot, Z, Zp, Zpp, rhs, qup, phirhs);
    for (size_t i = 0; i < tot; ++i) {
        rhs[2 * tot + 2 * i] = phirhs[i];
        rhs[2 * tot + 2 * i + 1] = 0.0;
    }
    if (diag) {
        if (diag->mu) memcpy(diag->mu, mu, sizeof(double) * tot);
        if (diag->Zp) memcpy(diag->Zp, Zp, sizeof(double) * 2 * tot);
        if (diag->Zpp) memcpy(diag->Zpp, Zpp, sizeof(double) * 2 * tot);
        if (diag->Phip) memcpy(diag->Phip, Phip, sizeof(double) * 2 * tot);
        if (diag->mup) memcpy(diag->mup, mup, sizeof(double) * 2 * tot);
        if (diag->qup) memcpy(diag->qup, qup, sizeof(double) * 2 * tot);
    }
    free(Zp); free(Zpp); free(Phip); free(mu); free(mucx); free(mup); free(qup); free(phirhs);
    free(M); free(piv);
    return info;
}

/* -------------------------------------------------------------------- RK4 */
typedef int (*sfo_rhs_fn)(void *ctx, const double *y, double *k);

/* AutonomousRungeKuttaStepperBase::runStep (AutonomousRungeKuttaStepper.cuh:123-307) with the
 * complex Zaxpy stages (:344-374).  y holds m complex values (2m doubles).  One step:
 *   k1 = f(y0); y1 = y0 + (dt/2) k1; k2 = f(y1); y2 = y0 + (dt/2) k2;
 *   k3 = f(y2); y3 = y0 + dt k3;     k4 = f(y3);
 *   K = ((k1 + 2 k2) + 2 k3) + k4  (add_k_vectors, utilities.cuh:78-83);  y0 += (dt/6) K. */
int sfo_rk4_generic(sfo_rhs_fn f, void *ctx, size_t m, double dt, int steps, double *y0, double *traj)
{
    const size_t len = 2 * m;
    const double half = dt * 0.5, sixth = dt / 6.0;   /* setTimeStep, :31-36 */
    double *k1 = (double *)malloc(sizeof(double) * len);
    double *k2 = (double *)malloc(sizeof(double) * len);
    double *k3 = (double *)malloc(sizeof(double) * len);
    double *k4 = (double *)malloc(sizeof(double) * len);
    double *ys = (double *)malloc(sizeof(double) * len);
    int info = 0, r;
    for (int s = 0; s < steps; ++s) {
        if ((r = f(ctx, y0, k1)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) ys[i] = y0[i] + half * k1[i];
        if ((r = f(ctx, ys, k2)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) ys[i] = y0[i] + half * k2[i];
        if ((r = f(ctx, ys, k3)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) ys[i] = y0[i] + dt * k3[i];
        if ((r = f(ctx, ys, k4)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) {
            double kk = k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i];
            y0[i] = y0[i] + sixth * kk;
        }
        if (traj) memcpy(traj + (size_t)s * len, y0, sizeof(double) * len);
    }
    free(k1); free(k2); free(k3); free(k4); free(ys);
    return info;
}

typedef struct { const sfo_props *p; int n, batch; } rhs_ctx;

static int rhs_adapter(void *ctx, const double *y, double *k)
{
    rhs_ctx *c = (rhs_ctx *)ctx;
    return sfo_rhs(c->p, c->n, c->batch, y, k, NULL);
}

int sfo_rk4(const sfo_props *p, int n, int batch, double dt, int steps, double *state, double *traj)
{
    rhs_ctx c = {p, n, batch};
    return sfo_rk4_generic(rhs_adapter, &c, (size_t)2 * n * batch, dt, steps, state, traj);
}

/* dx/dt = -y, dy/dt = x on each complex entry z = x + i y, i.e. dz/dt = i z (the flip_x_y
 * oscillator of the reference's RK4 test, ODESolverTests.cuh:14-23). */
static int rotation_rhs(void *ctx, const double *y, double *k)
{
    size_t m = *(size_t *)ctx;
    for (size_t i = 0; i < m; ++i) { k[2 * i] = -y[2 * i + 1]; k[2 * i + 1] = y[2 * i]; }
    return 0;
}

int sfo_rk4_rotation(size_t m, double dt, int steps, double *y)
{
    return sfo_rk4_generic(rotation_rhs, &m, m, dt, steps, y, NULL);
}

/* ------------------------------------------------------- optical drive */
/* Nondimensional optomechanical drive (OptomechanicalVariables after
 * adimensionalizeOptomechanicalVariables, Export.cu:1248-1270).  strength is
 * LightIntensity::get_current_intensity_drive_strength (LightIntensity.cuh:31-34). */
typedef struct sfo_drive {
    int augmented;    /* 1: state [Z, Phi, D], dD/dt = beta I - D / tau, dPhi/dt += D */
    int pad_;
    double damping;   /* DampingStrength */
    double strength;
    double detuning, gamma, G, max_intensity, x0, sigma, beta, tau;
} sfo_drive;

/* I = gamma^2 Imax / 4 / ((detuning - G y)^2 + (gamma/2)^2) * exp(-(x - x0)^2 / (2 sigma^2)) */
double sfo_intensity(const sfo_drive *d, double y, double x)
{
    const double df = d->detuning - d->G * y;
    const double prof = exp(-pow(x - d->x0, 2.0) / (2.0 * pow(d->sigma, 2.0)));
    return 0.25 * pow(d->gamma, 2.0) * d->max_intensity / (pow(df, 2.0) + pow(d->gamma / 2.0, 2.0)) * prof;
}

/* state / rhs: complex [Z(T), Phi(T)(, D(T))] (T = n batch nodes), Phi and D with imaginary part 0.
 * dPhi/dt = ((Bernoulli + damping v) + strength I) + D, in the reference's order of additions. */
int sfo_rhs_driven(const sfo_props *p, const sfo_drive *d, int n, int batch, const double *state, double *rhs)
{
    const size_t tot = (size_t)n * batch;
    const int info = sfo_rhs(p, n, batch, state, rhs, NULL);
    for (size_t i = 0; i < tot; ++i) {
        const double x = state[2 * i], y = state[2 * i + 1], v = rhs[2 * i + 1];
        const double inten = sfo_intensity(d, y, x);
        double r = rhs[2 * tot + 2 * i];
        r += d->damping * v;
        r += d->strength * inten;
        if (d->augmented) {
            const double D = state[4 * tot + 2 * i];
            r += D;
            rhs[4 * tot + 2 * i] = d->beta * inten - 1.0 / d->tau * D;
            rhs[4 * tot + 2 * i + 1] = 0.0;
        }
        rhs[2 * tot + 2 * i] = r;
    }
    return info;
}

typedef struct { const sfo_props *p; const sfo_drive *d; int n, batch; } drive_ctx;

static int driven_adapter(void *ctx, const double *y, double *k)
{
    drive_ctx *c = (drive_ctx *)ctx;
    return sfo_rhs_driven(c->p, c->d, c->n, c->batch, y, k);
}

int sfo_rk4_driven(const sfo_props *p, const sfo_drive *d, int n, int batch, double dt, int steps, double *state,
                   double *traj)
{
    drive_ctx c = {p, d, n, batch};
    return sfo_rk4_generic(driven_adapter, &c, (size_t)(d->augmented ? 3 : 2) * n * batch, dt, steps, state, traj);
}

/* Time-dependent drive (HeliumWithOptomechanicalDrivingProblem + TimedBoundaryIntegrator +
 * RungeKuttaStepper of RK4_Time_Dependent.cuh:170-412, add_optical_field_drive_terms createM.cuh:119-136,
 * DelayedIntensityTerm.cuh:14-38).  The stages of the step from t_n run at t_n, t_n + h/2, t_n + h/2
 * and t_n + h (h/2 = h * 0.5, t_n accumulated as currentTime += h from t0).  At a stage of time ts the
 * delayed intensity of node i is
 *     Ds = I_i                                       if ts == prev
 *     Ds = a Dsav_i + beta tau (1 - a) I_i,  a = exp(-(ts - prev) / tau)   otherwise,
 * with I_i the intensity at the stage state; the first stage of every step then stores Dsav = Ds and
 * prev = t_n (prev starts at t0: setStartingTime).  dPhi/dt = ((Bernoulli + damping v) + beta Ds)
 * + strength I.  Every node of a stage sees the same prev: the reference stores prev from every
 * thread of the first stage while other threads still read it, and this restatement takes the
 * order in which all reads come first. */
typedef struct {
    const sfo_props *p;
    const sfo_drive *d;
    int n, batch;
    double ts, prev;
    int save;
    double *dsav;
} timed_ctx;

static int timed_rhs(void *ctx, const double *y, double *k)
{
    timed_ctx *c = (timed_ctx *)ctx;
    const sfo_drive *d = c->d;
    const size_t tot = (size_t)c->n * c->batch;
    const int info = sfo_rhs(c->p, c->n, c->batch, y, k, NULL);
    const int same = c->ts == c->prev;
    const double a = same ? 0.0 : exp(-(c->ts - c->prev) / d->tau);
    const double bt = same ? 1.0 : d->beta * d->tau * (1.0 - a);
    for (size_t i = 0; i < tot; ++i) {
        const double x = y[2 * i], yy = y[2 * i + 1], v = k[2 * i + 1];
        const double inten = sfo_intensity(d, yy, x);
        const double D = same ? inten : a * c->dsav[i] + bt * inten;
        if (c->save) c->dsav[i] = D;
        double r = k[2 * tot + 2 * i];
        r += d->damping * v;
        r += d->beta * D;
        r += d->strength * inten;
        k[2 * tot + 2 * i] = r;
    }
    return info;
}

int sfo_rk4_timed(const sfo_props *p, const sfo_drive *d, int n, int batch, double dt, double t0, int steps,
                  double *state, double *dsav, double *traj)
{
    const size_t len = 4 * (size_t)n * batch;
    const double half = dt * 0.5, sixth = dt / 6.0;
    double *k1 = (double *)malloc(sizeof(double) * len), *k2 = (double *)malloc(sizeof(double) * len);
    double *k3 = (double *)malloc(sizeof(double) * len), *k4 = (double *)malloc(sizeof(double) * len);
    double *ys = (double *)malloc(sizeof(double) * len);
    timed_ctx c = {p, d, n, batch, t0, t0, 0, dsav};
    double t = t0;
    int info = 0, r;
    for (int s = 0; s < steps; ++s) {
        c.ts = t; c.save = 1;
        if ((r = timed_rhs(&c, state, k1)) && !info) info = r;
        c.prev = t; c.save = 0;
        for (size_t i = 0; i < len; ++i) ys[i] = state[i] + half * k1[i];
        c.ts = t + half;
        if ((r = timed_rhs(&c, ys, k2)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) ys[i] = state[i] + half * k2[i];
        if ((r = timed_rhs(&c, ys, k3)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) ys[i] = state[i] + dt * k3[i];
        c.ts = t + dt;
        if ((r = timed_rhs(&c, ys, k4)) && !info) info = r;
        for (size_t i = 0; i < len; ++i) {
            double kk = k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i];
            state[i] = state[i] + sixth * kk;
        }
        if (traj) memcpy(traj + (size_t)s * len, state, sizeof(double) * len);
        t += dt;
    }
    free(k1); free(k2); free(k3); free(k4); free(ys);
    return info;
}

/* ------------------------------------------- FD Jacobian and Gauss-Legendre */
/* Right-hand side of a real state [x(n), y(n), phi(n)] -> [vx(n), vy(n), dphi(n)]
 * (RealBoundaryItegralCalculator, RealBoundaryIntegralCalculator.cuh:4-40). */
typedef int (*sfo_real_fn)(void *ctx, const double *y, double *f);

typedef struct { const sfo_props *p; int n; } real_ctx;

static int real_rhs(void *ctx, const double *s, double *f)
{
    real_ctx *c = (real_ctx *)ctx;
    const int n = c->n;
    double *st = (double *)calloc(4 * (size_t)n + 1, sizeof(double));
    double *rh = (double *)calloc(4 * (size_t)n + 1, sizeof(double));
    for (int i = 0; i < n; ++i) {
        st[2 * i] = s[i];
        st[2 * i + 1] = s[n + i];
        st[2 * n + 2 * i] = s[2 * n + i];
        st[2 * n + 2 * i + 1] = 0.0;
    }
    int info = sfo_rhs(c->p, n, 1, st, rh, NULL);
    for (int i = 0; i < n; ++i) {
        f[i] = rh[2 * i];
        f[n + i] = rh[2 * i + 1];
        f[2 * n + i] = rh[2 * n + 2 * i];
    }
    free(st);
    free(rh);
    return info;
}

/* J[k m + r] = (f_r(s + eps e_k) - f_r(s - eps e_k)) / (2 eps), m = dim (column-major: column k is the
 * perturbed variable, createJacobianMatrixFromPerturbedRhs, JacobianCalculator.cuh:68-104). */
static int jacobian_generic(sfo_real_fn f, void *ctx, int m, const double *s, double eps, double *J)
{
    double *sp = (double *)malloc(sizeof(double) * m);
    double *fp = (double *)malloc(sizeof(double) * m);
    double *fm = (double *)malloc(sizeof(double) * m);
    int info = 0, r;
    for (int k = 0; k < m; ++k) {
        memcpy(sp, s, sizeof(double) * m);
        sp[k] = s[k] + eps;
        if ((r = f(ctx, sp, fp)) && !info) info = r;
        sp[k] = s[k] - eps;
        if ((r = f(ctx, sp, fm)) && !info) info = r;
        for (int q = 0; q < m; ++q) J[(size_t)k * m + q] = (fp[q] - fm[q]) / (2.0 * eps);
    }
    free(sp);
    free(fp);
    free(fm);
    return info;
}

int sfo_jacobian(const sfo_props *p, int n, const double *s, double eps, double *J)
{
    real_ctx c = {p, n};
    return jacobian_generic(real_rhs, &c, 3 * n, s, eps, J);
}

/* Gauss-Legendre two-stage coefficients (GLCoefficients.hpp): A = [[1/4, 1/4 - sqrt3/6],
 * [1/4 + sqrt3/6, 1/4]], b = [1/2, 1/2]. */
typedef struct {
    double tol, armijo_c, backtrack, min_alpha;
    int max_newton, max_halves;
} gl2_opts;

static double gl2_residual(sfo_real_fn f, void *ctx, int m, const double *y, double h, const double *k, double *R,
                           double *ys, double *fs)
{
    const double s3 = 1.7320508075688772935;
    const double A11 = 0.25, A12 = 0.25 - s3 / 6.0, A21 = 0.25 + s3 / 6.0, A22 = 0.25;
    const double *k1 = k, *k2 = k + m;
    for (int i = 0; i < m; ++i) ys[i] = y[i] + h * (A11 * k1[i] + A12 * k2[i]);
    f(ctx, ys, fs);
    for (int i = 0; i < m; ++i) R[i] = k1[i] - fs[i];
    for (int i = 0; i < m; ++i) ys[i] = y[i] + h * (A21 * k1[i] + A22 * k2[i]);
    f(ctx, ys, fs);
    for (int i = 0; i < m; ++i) R[m + i] = k2[i] - fs[i];
    double ss = 0.0;
    for (int i = 0; i < 2 * m; ++i) ss += R[i] * R[i];
    return ss;
}

/* One step of size h from y into yn.  Newton on R(k) = [k1 - f(y + h(A11 k1 + A12 k2)),
 * k2 - f(y + h(A21 k1 + A22 k2))] with the matrix [[I - h A11 J1, -h A12 J1], [-h A21 J2, I - h A22 J2]]
 * (J_i the Jacobian at stage state i), Armijo backtracking on |R|^2 / 2.  Returns 1 if converged. */
static int gl2_step(sfo_real_fn f, void *ctx, int m, const double *y, double h, const gl2_opts *o, double eps,
                    double *yn)
{
    const double s3 = 1.7320508075688772935;
    const double A[2][2] = {{0.25, 0.25 - s3 / 6.0}, {0.25 + s3 / 6.0, 0.25}};
    const int m2 = 2 * m;
    double *k = (double *)malloc(sizeof(double) * m2), *kt = (double *)malloc(sizeof(double) * m2);
    double *R = (double *)malloc(sizeof(double) * m2), *dk = (double *)malloc(sizeof(double) * m2);
    double *ys = (double *)malloc(sizeof(double) * m), *fs = (double *)malloc(sizeof(double) * m);
    double *J1 = (double *)malloc(sizeof(double) * m * m), *J2 = (double *)malloc(sizeof(double) * m * m);
    double *Mx = (double *)malloc(sizeof(double) * m2 * m2);
    int *piv = (int *)malloc(sizeof(int) * m2);
    f(ctx, y, k);
    memcpy(k + m, k, sizeof(double) * m);
    int ok = 0;
    for (int it = 0; it < o->max_newton; ++it) {
        double ss = gl2_residual(f, ctx, m, y, h, k, R, ys, fs);
        double kk = 0.0;
        for (int i = 0; i < m2; ++i) kk += k[i] * k[i];
        if (sqrt(ss) <= o->tol * (1.0 + kk)) { ok = 1; break; }
        for (int st = 0; st < 2; ++st) {   /* stage states and their Jacobians */
            for (int i = 0; i < m; ++i) ys[i] = y[i] + h * (A[st][0] * k[i] + A[st][1] * k[m + i]);
            jacobian_generic(f, ctx, m, ys, eps, st ? J2 : J1);
        }
        for (int r = 0; r < m2; ++r) {   /* row-major Newton matrix */
            const int bi = r / m, rr = r % m;
            const double *J = bi ? J2 : J1;
            for (int c = 0; c < m2; ++c) {
                const int bj = c / m, cc = c % m;
                Mx[(size_t)r * m2 + c] = (r == c ? 1.0 : 0.0) - h * A[bi][bj] * J[(size_t)cc * m + rr];
            }
        }
        for (int i = 0; i < m2; ++i) dk[i] = -R[i];
        sfo_lu_factor(m2, Mx, piv);
        sfo_lu_solve(m2, Mx, piv, dk);
        double alpha = 1.0, phi = 0.5 * ss;
        for (;;) {
            for (int i = 0; i < m2; ++i) kt[i] = k[i] + alpha * dk[i];
            double sst = gl2_residual(f, ctx, m, y, h, kt, R, ys, fs);
            if (0.5 * sst <= phi - o->armijo_c * alpha * ss) { memcpy(k, kt, sizeof(double) * m2); break; }
            alpha *= o->backtrack;
            if (alpha < o->min_alpha) goto done;
        }
    }
    if (!ok) {   /* the last Newton update may have converged */
        double ss = gl2_residual(f, ctx, m, y, h, k, R, ys, fs), kk = 0.0;
        for (int i = 0; i < m2; ++i) kk += k[i] * k[i];
        ok = sqrt(ss) <= o->tol * (1.0 + kk);
    }
done:
    for (int i = 0; i < m; ++i) yn[i] = y[i] + h * (0.5 * k[i] + 0.5 * k[m + i]);
    free(k); free(kt); free(R); free(dk); free(ys); free(fs); free(J1); free(J2); free(Mx); free(piv);
    return ok;
}

/* runEvolution (GaussLegendre.cuh:190-290): steps of at most h until t1, halving a step that does
 * not converge (up to max_halves times) and keeping the reduced size afterwards.  traj receives the
 * initial state and every step's state (max_frames of them at most); returns the frame count, or
 * -1 when a step fails at every size. */
static int gl2_run(sfo_real_fn f, void *ctx, int m, double *y, double t0, double t1, double h, const gl2_opts *o,
                   double eps, double *traj, double *times, int max_frames)
{
    double *yn = (double *)malloc(sizeof(double) * m);
    int frames = 0;
    if (traj && frames < max_frames) { memcpy(traj, y, sizeof(double) * m); times[0] = t0; }
    frames = 1;
    double t = t0;
    while (t < t1) {
        double heff = h < t1 - t ? h : t1 - t, ht = heff;
        int ok = 0;
        for (int i = 0; i <= o->max_halves; ++i) {
            if ((ok = gl2_step(f, ctx, m, y, ht, o, eps, yn))) break;
            ht *= 0.5;
        }
        if (!ok) { free(yn); return -1; }
        memcpy(y, yn, sizeof(double) * m);
        t += ht;
        h = ht;
        if (traj && frames < max_frames) {
            memcpy(traj + (size_t)frames * m, y, sizeof(double) * m);
            times[frames] = t;
        }
        ++frames;
    }
    free(yn);
    return frames;
}

int sfo_gl2(const sfo_props *p, int n, double *s, double t0, double t1, double h, double tol, int max_newton,
            double eps, double *traj, double *times, int max_frames)
{
    real_ctx c = {p, n};
    gl2_opts o = {tol, 1e-4, 0.5, 1e-6, max_newton, 6};
    return gl2_run(real_rhs, &c, 3 * n, s, t0, t1, h, &o, eps, traj, times, max_frames);
}

/* dx/dt = -y, dy/dt = x on m/2 pairs (the rotation of the RK4 test), for pinning the integrator */
static int rotation_real(void *ctx, const double *y, double *f)
{
    const int h = *(int *)ctx;
    for (int i = 0; i < h; ++i) { f[i] = -y[h + i]; f[h + i] = y[i]; }
    return 0;
}

int sfo_gl2_rotation(int pairs, double *y, double t1, double h, double tol)
{
    gl2_opts o = {tol, 1e-4, 0.5, 1e-6, 20, 6};
    return gl2_run(rotation_real, &pairs, 2 * pairs, y, 0.0, t1, h, &o, 1e-6, NULL, NULL, 0);
}

/* ------------------------------------------------------------- energies */
/* EnergyContainer::CalculateEnergy + VolumeFlux (Energies.cuh:56-204) for one surface of n nodes:
 * out[0] kinetic   (1/4pi) sum_k (phi_k + U(1+rho)x_k/2)(-Y'_k u_k + X'_k v_k)
 *                          - (U/2)((u_k + rho u_k) X'_k + (v_k + rho v_k) Y'_k + U(1-rho)X'_k/2) y_k
 *                  (KineticEnergy passes the lower velocity as the upper one too, Energies.cuh:256)
 * out[1] potential water: (1+rho)/4pi sum_k y_k^2 X'_k (GravitationalEnergy);
 *                  helium: h^2/6 sum_k ((1 + y_k/h)^-2 - 1) (VanDerWaalsEnergy); both times g
 *                  (g = 1 is the reference, which has no g)
 * out[2] surface   kappa/2pi (sum_k |Z'_k| - 2pi)
 * out[3] flux      1/2pi sum_k (v_k X'_k + u_k Y'_k)
 * with Z = x + i y, Z' = X' + i Y', q = u + i v (the lower velocity after conjugation). */
void sfo_energies(const sfo_props *p, int n, const double *Z, const double *Phi, const double *Zp,
                  const double *q, double *out)
{
    const double U = p->U, rho = p->rho, h = p->depth;
    double kin = 0, pot = 0, surf = 0, flux = 0;
    for (int k = 0; k < n; ++k) {
        const double x = Z[2 * k], y = Z[2 * k + 1], X = Zp[2 * k], Y = Zp[2 * k + 1];
        const double u = q[2 * k], v = q[2 * k + 1];
        kin += (Phi[k] + 0.5 * U * (1.0 + rho) * x) * (-Y * u + X * v) -
               0.5 * U * ((u + rho * u) * X + (v + rho * v) * Y + 0.5 * U * (1.0 - rho) * X) * y;
        if (p->kind == 0) pot += y * y * X;
        else pot += 1.0 / pow(1.0 + y / h, 2) - 1.0;
        surf += sqrt(X * X + Y * Y);
        flux += v * X + u * Y;
    }
    out[0] = kin * 0.25 / PI_D;
    out[1] = p->g * (p->kind == 0 ? pot * 0.25 * (1.0 + rho) / PI_D : pot * h * h / 6.0);
    out[2] = (surf - 2.0 * PI_D) * p->kappa / (2.0 * PI_D);
    out[3] = flux * 0.5 / PI_D;
}

/* -------------------------------------------------- SI to base units */
/* adimensionalizeProperties, Export.cu:1222-1246 (rhoHelium = 150, alpha = 3.5e-24). */
void sfo_adimensionalize(double L, double rho, double kappa, double depth, double *out)
{
    const double alpha = 3.5e-24, rho_he = 150.0;
    double base_length = L / (2.0 * PI_D);
    double base_acc = 3 * alpha / pow(depth, 4);
    double base_time = sqrt(base_length / base_acc);
    double stf = rho_he * base_length * base_length * base_length / (base_time * base_time);
    out[0] = rho / rho_he;
    out[1] = kappa / stf;
    out[2] = depth / base_length;
    out[3] = base_length;
    out[4] = base_time;
}
