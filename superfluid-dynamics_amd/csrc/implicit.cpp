// implicit.cpp -- central-difference Jacobian of the right-hand side and the two-stage
// Gauss-Legendre integrator built on it (host orchestration of the gfx950 stepper's kernels).
//
// Jacobian (JacobianCalculator::calculateJacobian, CuSuperHelium/JacobianCalculator.cuh:223-284):
// the 3N variables of a real state [x, y, phi] are perturbed by +-eps one at a time and the 6N
// perturbed surfaces evaluated as uniform batches of the FP64 LU path (the reference's
// getrfBatched), at most kJacBudget bytes of dipole matrices per batch;
// J[k 3N + r] = (f_r(s + eps e_k) - f_r(s - eps e_k)) / (2 eps), column-major as the reference stores
// it (createJacobianMatrixFromPerturbedRhs, JacobianCalculator.cuh:68-104).
//
// Gauss-Legendre (GaussLegendre2::runEvolution / gaussLegendreS2Step, GaussLegendre.cuh:190-560,
// coefficients GLCoefficients.hpp): stage slopes k1, k2 solve
//   R1 = k1 - f(y + h (A11 k1 + A12 k2)) = 0,  R2 = k2 - f(y + h (A21 k1 + A22 k2)) = 0
// by Newton with the matrix [[I - h A11 J1, -h A12 J1], [-h A21 J2, I - h A22 J2]] (J_i the Jacobian at
// stage state i; solved by the device LU with its GEPP fallback), Armijo backtracking on |R|^2 / 2,
// an optional frozen-Jacobian fallback, and step halving when a step does not converge; then
// y += h (k1 + k2) / 2.  Converged when |R| <= tol (1 + |k|^2).  The reference builds its Newton
// matrix from the transposed Jacobian (fillMMatrix reads the column-major J row-major) and measures
// only R1; either changes the iteration path, not the solution the iteration converges to.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "solver.hpp"

namespace sfh {

namespace {

constexpr double kJacBudget = 4.0 * (1u << 30);   // bytes of dipole matrices per perturbation batch

// Right-hand sides of `count` real states [x, y, phi] (count x 3n, contiguous) -> count x 3n.
class BatchRhs {
  public:
    BatchRhs(const Phys& ph, int n, int count) : n_(n), count_(count) {
        Options o;
        o.solver = kSolveLU;
        solver_ = std::make_unique<Solver>(ph, std::vector<int>(count, n), o);
        x_.resize((size_t)n * count);
        y_.resize(x_.size());
        p_.resize(x_.size());
        vx_.resize(x_.size());
        vy_.resize(x_.size());
        dp_.resize(x_.size());
    }
    int count() const { return count_; }
    void eval(const double* states, double* rhs, int used) {
        const size_t n = n_;
        for (int b = 0; b < count_; ++b) {
            const double* s = states + 3 * n * std::min(b, used - 1);   // unused slots repeat the last state
            std::memcpy(&x_[b * n], s, n * sizeof(double));
            std::memcpy(&y_[b * n], s + n, n * sizeof(double));
            std::memcpy(&p_[b * n], s + 2 * n, n * sizeof(double));
        }
        solver_->set_state(x_.data(), y_.data(), p_.data());
        solver_->rhs(vx_.data(), vy_.data(), dp_.data());
        for (int b = 0; b < used; ++b) {
            double* r = rhs + 3 * n * b;
            std::memcpy(r, &vx_[b * n], n * sizeof(double));
            std::memcpy(r + n, &vy_[b * n], n * sizeof(double));
            std::memcpy(r + 2 * n, &dp_[b * n], n * sizeof(double));
        }
    }

  private:
    int n_, count_;
    std::unique_ptr<Solver> solver_;
    std::vector<double> x_, y_, p_, vx_, vy_, dp_;
};

}  // namespace

class FdJacobian {
  public:
    FdJacobian(const Phys& ph, int n) : n_(n) {
        const int m = 3 * n;
        const double per = 8.0 * n * (double)n;
        const int chunk = (int)std::max(1.0, std::min((double)(2 * m), std::floor(kJacBudget / per)));
        batch_ = std::make_unique<BatchRhs>(ph, n, chunk);
    }
    // jac: column-major 3n x 3n
    void compute(const double* s, double eps, double* jac) {
        if (!(eps > 0.0) || !std::isfinite(eps)) throw ArgError("epsilon must be positive");
        const int m = 3 * n_, B = batch_->count();
        std::vector<double> states((size_t)B * m), out((size_t)B * m);
        std::vector<double> fp((size_t)m * m), fm((size_t)m * m);   // f(s + eps e_k), f(s - eps e_k), by k
        for (int j0 = 0; j0 < 2 * m; j0 += B) {
            const int used = std::min(B, 2 * m - j0);
            for (int b = 0; b < used; ++b) {
                const int j = j0 + b, k = j % m;
                double* st = &states[(size_t)b * m];
                std::memcpy(st, s, m * sizeof(double));
                st[k] = j < m ? s[k] + eps : s[k] - eps;
            }
            batch_->eval(states.data(), out.data(), used);
            for (int b = 0; b < used; ++b) {
                const int j = j0 + b;
                double* dst = (j < m ? fp.data() : fm.data()) + (size_t)(j % m) * m;
                std::memcpy(dst, &out[(size_t)b * m], m * sizeof(double));
            }
        }
        for (size_t i = 0; i < (size_t)m * m; ++i) jac[i] = (fp[i] - fm[i]) / (2.0 * eps);
    }

  private:
    int n_;
    std::unique_ptr<BatchRhs> batch_;
};

void fd_jacobian(const Phys& ph, int n, const double* state, double eps, double* jac) {
    if (!supported_n(n) || n > 2048) throw ArgError("unsupported node count " + std::to_string(n));
    FdJacobian J(ph, n);
    J.compute(state, eps, jac);
}

namespace {

struct GLCoeffs {
    double A[2][2];
    double b[2];
    GLCoeffs() {
        const double s3 = 1.7320508075688772935;
        A[0][0] = 0.25;
        A[0][1] = 0.25 - s3 / 6.0;
        A[1][0] = 0.25 + s3 / 6.0;
        A[1][1] = 0.25;
        b[0] = b[1] = 0.5;
    }
};

class GaussLegendre2 {
  public:
    GaussLegendre2(const Phys& ph, int n, const Gl2Options& o) : m_(3 * n), o_(o), pair_(ph, n, 2), jac_(ph, n) {}

    // One step of size h from y into yn; returns whether the Newton iteration converged.
    bool step(const double* y, double h, double* yn) {
        const int m = m_, m2 = 2 * m;
        std::vector<double> k(m2), kt(m2), R(m2), dk(m2), J1((size_t)m * m), J2((size_t)m * m), Jf;
        std::vector<double> Mx((size_t)m2 * m2);
        f_single(y, k.data());
        std::memcpy(&k[m], k.data(), m * sizeof(double));
        bool frozen = false, ok = false;
        for (int it = 0; it < o_.max_newton; ++it) {
            const double ss = residual(y, h, k.data(), R.data());
            if (converged(ss, k)) {
                ok = true;
                break;
            }
            if (!frozen) {
                std::vector<double> ys(m);
                for (int st = 0; st < 2; ++st) {
                    stage_state(y, h, k.data(), st, ys.data());
                    jac_.compute(ys.data(), o_.eps, st ? J2.data() : J1.data());
                }
            }
            for (int r = 0; r < m2; ++r) {
                const int bi = r / m, rr = r % m;
                const double* J = bi ? J2.data() : J1.data();
                for (int c = 0; c < m2; ++c) {
                    const int bj = c / m, cc = c % m;
                    Mx[(size_t)r * m2 + c] = (r == c ? 1.0 : 0.0) - h * co_.A[bi][bj] * J[(size_t)cc * m + rr];
                }
            }
            for (int i = 0; i < m2; ++i) R[i] = -R[i];
            int inter = 0;
            debug_lu_solve(m2, Mx.data(), R.data(), dk.data(), &inter);
            double alpha = 1.0;
            const double phi = 0.5 * ss;
            bool accepted = false;
            while (alpha >= o_.min_alpha) {
                for (int i = 0; i < m2; ++i) kt[i] = k[i] + alpha * dk[i];
                std::vector<double> Rt(m2);
                const double sst = residual(y, h, kt.data(), Rt.data());
                if (0.5 * sst <= phi - o_.armijo_c * alpha * ss) {
                    k.swap(kt);
                    accepted = true;
                    break;
                }
                alpha *= o_.backtrack;
            }
            if (!accepted) {
                if (!frozen && o_.simplified_fallback) {   // the Jacobian at y for both stages from now on
                    frozen = true;
                    jac_.compute(y, o_.eps, J1.data());
                    J2 = J1;
                    continue;
                }
                break;
            }
        }
        if (!ok) ok = converged(residual(y, h, k.data(), R.data()), k);
        for (int i = 0; i < m; ++i) yn[i] = y[i] + h * (co_.b[0] * k[i] + co_.b[1] * k[m + i]);
        return ok;
    }

  private:
    int m_;
    Gl2Options o_;
    GLCoeffs co_;
    BatchRhs pair_;   // the two stage states in one launch sequence
    FdJacobian jac_;

    void f_single(const double* y, double* f) { pair_.eval(y, f, 1); }
    void stage_state(const double* y, double h, const double* k, int st, double* ys) const {
        for (int i = 0; i < m_; ++i) ys[i] = y[i] + h * (co_.A[st][0] * k[i] + co_.A[st][1] * k[m_ + i]);
    }
    double residual(const double* y, double h, const double* k, double* R) {
        std::vector<double> ys(2 * (size_t)m_), fs(2 * (size_t)m_);
        stage_state(y, h, k, 0, ys.data());
        stage_state(y, h, k, 1, ys.data() + m_);
        pair_.eval(ys.data(), fs.data(), 2);
        double ss = 0.0;
        for (int i = 0; i < 2 * m_; ++i) {
            R[i] = k[i] - fs[i];
            ss += R[i] * R[i];
        }
        return ss;
    }
    bool converged(double ss, const std::vector<double>& k) const {
        double kk = 0.0;
        for (double v : k) kk += v * v;
        return std::sqrt(ss) <= o_.tol * (1.0 + kk);
    }
};

}  // namespace

int gl2_integrate(const Phys& ph, int n, double* state, double t0, double t1, double h, const Gl2Options& o,
                  std::vector<double>* traj, std::vector<double>* times) {
    if (!supported_n(n) || n > 1024) throw ArgError("unsupported node count " + std::to_string(n));
    if (!(h > 0.0) || !std::isfinite(h)) throw ArgError("step size must be positive");
    if (!std::isfinite(t0) || !std::isfinite(t1) || t1 < t0) throw ArgError("need finite t0 <= t1");
    const int m = 3 * n;
    GaussLegendre2 gl(ph, n, o);
    std::vector<double> yn(m);
    if (traj) {
        traj->assign(state, state + m);
        times->assign(1, t0);
    }
    const double hmin = (t1 - t0) / std::pow(2.0, 20);
    double t = t0;
    int steps = 0;
    while (t < t1) {
        const double heff = std::min(h, t1 - t);
        double ht = heff;
        bool ok = false;
        for (int i = 0; i <= o.max_halves; ++i) {
            if ((ok = gl.step(state, ht, yn.data()))) break;
            if (ht <= hmin) break;
            ht *= 0.5;
        }
        if (!ok) throw HipError("Gauss-Legendre step did not converge at t = " + std::to_string(t));
        std::memcpy(state, yn.data(), m * sizeof(double));
        t += ht;
        h = ht;   // the reduced step size is kept (GaussLegendre.cuh:280)
        ++steps;
        if (traj) {
            traj->insert(traj->end(), state, state + m);
            times->push_back(t);
        }
    }
    return steps;
}

}  // namespace sfh
