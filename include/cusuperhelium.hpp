// cusuperhelium.hpp -- C++ plugin API of the gfx950 Roberts RK4 stepper, header-only over the C ABI
// of cusuperhelium.h.
//
// Mirrors the types a C++ caller of the reference builds a simulation from (paths relative to
// CuSuperHelium/CuSuperHelium/):
//   ProblemProperties                     ProblemProperties.hpp:5-32
//   RK4Options, OdeSolverResult           RK4Options.h, OdeSolver.h
//   AutonomousProblem<T, N>               AutonomousProblem.h:9-28 (run(state, rhs), setStream)
//   WaterBoundaryProblem<N, B>            WaterBoundaryProblem.cuh:7-40
//   HeliumBoundaryProblem<N, B>           HeliumBoundaryProblem.cuh:7-46
//   BaseBoundaryIntegralCalculator<N, B>  BaseBoundaryIntegrator.cuh:10-85 (an AutonomousProblem of
//                                         2 N B complex values: state = [Z(N B), Phi(N B)])
//   TrajectoryLogger<T, N>                TrajectoryLogger.cuh:8-73
//   AutonomousRungeKuttaStepper<T, N>     AutonomousRungeKuttaStepper.cuh:24-89,309-329,417-437
// with the same template parameters, constructor arguments and member functions, so that code
// written against those headers compiles against this one (namespace cusuperhelium; std_complex is
// std::complex<double>, layout-compatible with the reference's cuda::std::complex<double>).
//
// Differences, by design:
//   * the boundary problem objects only select the physics; the kernels behind them are this
//     library's (no CreateMMatrix / CalculateVelocities virtuals: the matrix and the velocity
//     matrices are never materialised);
//   * the stepper advances the calculator's fused RK4 step graph, so it takes a
//     BaseBoundaryIntegralCalculator rather than an arbitrary AutonomousProblem;
//   * errors throw cusuperhelium::SolverError (the reference prints or exits);
//   * streams are passed as void* (a hipStream_t) so that this header needs no HIP headers.
#pragma once
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cusuperhelium.h"

namespace cusuperhelium {

using std_complex = std::complex<double>;
static_assert(sizeof(std_complex) == 2 * sizeof(double), "complex128 layout");

struct SolverError : std::runtime_error {
    int code;
    SolverError(int c, const std::string& where)
        : std::runtime_error(where + " failed with code " + std::to_string(c) + ": " + sfh_last_error()), code(c) {}
};

inline void check(int code, const char* where) {
    if (code != 0) throw SolverError(code, where);
}

// ProblemProperties.hpp:5-32 (same fields and defaults)
struct ProblemProperties {
    double L = 1.0;
    double rho = 1.0;
    double U = 0.0;
    double kappa = 0.0;
    double depth = 1.0;
    double initial_amplitude = 1.0;
    double y_min = 0.0;
    double y_max = 0.0;
    bool use_expansions = false;
    int expansion_order = 1;
    bool infinite_depth = false;
    double base_length = 1.0;
    double base_time = 1.0;
    double base_energy = 1.0;
    double base_acceleration = 1.0;
};

// RK4Options.h
struct RK4Options {
    double initial_timestep = 1e-2;
    bool returnTrajectory = true;
};

// OdeSolver.h
enum class OdeSolverResult { ReachedEndTime, StiffnessDetected };

// AutonomousProblem.h:9-28
template <typename T, int N>
class AutonomousProblem {
  public:
    virtual ~AutonomousProblem() {}
    virtual void run(T* initialState, T* rhs) = 0;
    virtual void setStream(void* stream) = 0;
};

// The physics of a boundary problem (which kernels the calculator runs); N and batchSize as in the
// reference's templates.
template <int N, size_t batchSize>
class BoundaryProblem {
  public:
    explicit BoundaryProblem(int kind, ProblemProperties& properties) : kind_(kind), properties_(properties) {}
    virtual ~BoundaryProblem() {}
    int kind() const { return kind_; }
    ProblemProperties& properties() const { return properties_; }

  private:
    int kind_;
    ProblemProperties& properties_;
};

// WaterBoundaryProblem.cuh:7-40: createMKernel, createVelocityMatrices, compute_rhs_phi_expression
template <int N, size_t batchSize>
class WaterBoundaryProblem : public BoundaryProblem<N, batchSize> {
  public:
    explicit WaterBoundaryProblem(ProblemProperties& properties) : BoundaryProblem<N, batchSize>(SFH_WATER, properties) {}
};

// HeliumBoundaryProblem.cuh:7-46: createFiniteDepthMKernel, createHeliumVelocityMatrices, the
// helium Bernoulli kernels (expansions > kappa != 0 > plain)
template <int N, size_t batchSize>
class HeliumBoundaryProblem : public BoundaryProblem<N, batchSize> {
  public:
    explicit HeliumBoundaryProblem(ProblemProperties& properties) : BoundaryProblem<N, batchSize>(SFH_HELIUM, properties) {}
};

// Solver selection for the calculator (not in the reference, which always uses getrf/getrs).
struct CalculatorOptions {
    int solver = SFH_SOLVE_LU;
    int precision = SFH_FP64;
    int max_iter = 32;
    double tol = 1e-12;
};

// BaseBoundaryIntegrator.cuh:10-85: run(state, rhs) is one right-hand side of the batchSize
// surfaces, state and rhs device arrays of 2 N batchSize complex values.
template <int N, size_t batchSize>
class BaseBoundaryIntegralCalculator : public AutonomousProblem<std_complex, 2 * N * (int)batchSize> {
  public:
    BaseBoundaryIntegralCalculator(ProblemProperties& problemProperties, BoundaryProblem<N, batchSize>& boundaryProblem,
                                   const CalculatorOptions& options = CalculatorOptions())
        : properties_(problemProperties) {
        sfh_props p;
        sfh_default_props(&p);
        p.kind = boundaryProblem.kind();
        p.rho = problemProperties.rho;
        p.kappa = problemProperties.kappa;
        p.depth = problemProperties.depth;
        p.U = problemProperties.U;
        p.infinite_depth = problemProperties.infinite_depth ? 1 : 0;
        p.use_expansions = problemProperties.use_expansions ? 1 : 0;
        p.expansion_order = problemProperties.expansion_order;
        sfh_options o;
        sfh_default_options(&o);
        o.solver = options.solver;
        o.precision = options.precision;
        o.max_iter = options.max_iter;
        o.tol = options.tol;
        std::vector<int> sizes(batchSize, N);
        check(sfh_create(&p, sizes.data(), (int)batchSize, &o, &h_), "sfh_create");
    }
    ~BaseBoundaryIntegralCalculator() override {
        if (h_) sfh_destroy(h_);
    }
    BaseBoundaryIntegralCalculator(const BaseBoundaryIntegralCalculator&) = delete;
    BaseBoundaryIntegralCalculator& operator=(const BaseBoundaryIntegralCalculator&) = delete;

    void run(std_complex* initialState, std_complex* rhs) override {
        check(sfh_rhs_device(h_, initialState, rhs), "sfh_rhs_device");
    }
    // The handle runs on its own stream and every call returns completed, so a caller's stream
    // needs no ordering with it.
    void setStream(void* stream) override { stream_ = stream; }

    // dipole strengths ("vorticities", getDevA) of the last right-hand side, host copy
    std::vector<double> getA() const {
        std::vector<double> a(N * batchSize);
        check(sfh_get_mu(h_, a.data()), "sfh_get_mu");
        return a;
    }
    sfh_handle handle() const { return h_; }

  private:
    ProblemProperties& properties_;
    sfh_handle h_ = nullptr;
    void* stream_ = nullptr;
};

// TrajectoryLogger.cuh:8-73: the states after each logged step and their times; copies to host
// arrays allocated with malloc (released by the caller with free, as the reference's exports do).
template <typename T, int N>
class TrajectoryLogger {
  public:
    virtual ~TrajectoryLogger() {}
    virtual int copyTimesToHost(double** times, size_t* count) {
        if (!enabled_) return -1;
        double* t = static_cast<double*>(std::malloc((times_.size() ? times_.size() : 1) * sizeof(double)));
        if (!t) return -1;
        if (!times_.empty()) std::memcpy(t, times_.data(), times_.size() * sizeof(double));
        *times = t;
        *count = times_.size();
        return 0;
    }
    virtual int copyStatesToHost(T** hostStates, size_t* countHost) {
        if (!enabled_) return -1;
        const size_t frames = states_.size() / N;
        T* h = static_cast<T*>(std::malloc((frames ? frames : 1) * N * sizeof(T)));
        if (!h) return -1;
        if (frames) std::memcpy(h, states_.data(), frames * N * sizeof(T));
        *hostStates = h;
        *countHost = frames;
        return 0;
    }
    bool isEnabled() const { return enabled_; }
    void setStream(void*) {}
    // frames in the reference's state layout: N = 2 nodes complex values [Z, Phi + 0i]
    void logFrame(double time, const T* frame) {
        if (!enabled_) return;
        times_.push_back(time);
        states_.insert(states_.end(), frame, frame + N);
    }

  private:
    bool enabled_ = true;
    std::vector<double> times_;
    std::vector<T> states_;
};

// AutonomousRungeKuttaStepper.cuh:24-89: classical RK4 over the calculator's fused step.
template <typename T, int N>
class AutonomousRungeKuttaStepper {
  public:
    template <int M, size_t B>
    AutonomousRungeKuttaStepper(BaseBoundaryIntegralCalculator<M, B>& problem, double tstep = 1e-2,
                                std::shared_ptr<TrajectoryLogger<T, N>> logger = nullptr)
        : h_(problem.handle()), logger_(std::move(logger)) {
        static_assert(N == 2 * M * (int)B, "the stepper's state holds 2 N batchSize complex values");
        setTimeStep(tstep);
    }
    void setTimeStep(double tstep) { dt_ = tstep; }
    void setOptions(const RK4Options& options) { setTimeStep(options.initial_timestep); }

    // initialize(devY0, onDevice), AutonomousRungeKuttaStepper.cuh:309-329: with onDevice the
    // caller's device array is the state and each runStep advances it in place; otherwise y0 is a
    // host array that is copied in.
    void initialize(T* y0, bool onDevice = false) {
        static_assert(sizeof(T) == sizeof(std_complex), "complex128 state");
        if (onDevice) {
            devY0_ = y0;
            check(sfh_set_state_device(h_, y0), "sfh_set_state_device");
        } else {
            devY0_ = nullptr;
            const int T_ = N / 2;
            std::vector<double> x(T_), y(T_), p(T_);
            const std_complex* c = reinterpret_cast<const std_complex*>(y0);
            for (int i = 0; i < T_; ++i) {
                x[i] = c[i].real();
                y[i] = c[i].imag();
                p[i] = c[T_ + i].real();
            }
            check(sfh_set_state(h_, x.data(), y.data(), p.data()), "sfh_set_state");
        }
        currentTime_ = 0.0;
    }
    // runStep(step): one RK4 step of the state in place
    void runStep(int /*step*/) {
        check(sfh_step(h_, dt_, 1), "sfh_step");
        if (devY0_) check(sfh_get_state_device(h_, devY0_), "sfh_get_state_device");
    }
    // runEvolution(t0, t1), AutonomousRungeKuttaStepper.cuh:417-437: floor((t1 - t0)/dt) steps,
    // each logged (when a logger is set) with currentTime advanced by dt per step
    OdeSolverResult runEvolution(double startTime, double endTime) {
        currentTime_ = startTime;
        const size_t steps = static_cast<size_t>((endTime - startTime) / dt_);
        if (logger_ && steps) {
            const int T_ = N / 2;
            std::vector<double> frames(3 * (size_t)T_ * steps), times(steps);
            for (size_t done = 0; done < steps;) {
                const int chunk = (int)std::min<size_t>(steps - done, 1u << 20);
                check(sfh_step_record(h_, dt_, chunk, 1, frames.data() + 3 * (size_t)T_ * done, times.data() + done),
                      "sfh_step_record");
                done += (size_t)chunk;
            }
            std::vector<std_complex> st(N);
            for (size_t k = 0; k < steps; ++k) {
                const double* f = frames.data() + 3 * (size_t)T_ * k;
                for (int i = 0; i < T_; ++i) {
                    st[i] = std_complex(f[i], f[T_ + i]);
                    st[T_ + i] = std_complex(f[2 * T_ + i], 0.0);
                }
                currentTime_ += dt_;
                logger_->logFrame(currentTime_, reinterpret_cast<const T*>(st.data()));
            }
        } else {
            for (size_t done = 0; done < steps;) {
                const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                check(sfh_step(h_, dt_, chunk), "sfh_step");
                done += (size_t)chunk;
            }
            for (size_t k = 0; k < steps; ++k) currentTime_ += dt_;
        }
        if (devY0_) check(sfh_get_state_device(h_, devY0_), "sfh_get_state_device");
        return OdeSolverResult::ReachedEndTime;
    }
    double currentTime() const { return currentTime_; }

  private:
    sfh_handle h_;
    std::shared_ptr<TrajectoryLogger<T, N>> logger_;
    T* devY0_ = nullptr;
    double dt_ = 1e-2;
    double currentTime_ = 0.0;
};

}  // namespace cusuperhelium
